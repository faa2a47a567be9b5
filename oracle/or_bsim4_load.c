/* or_bsim4_load.c -- oracle restatement of the BSIM4 loads (test
 * infrastructure only).
 * Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET_B4.C
 *   setupCapacitors_newDAE :12357-12606, setupCapacitors_oldDAE :12616-13031,
 *   auxChargeCalculations :12216-12347, setupFVectorVars :13525-13838,
 *   registerLIDs :4996-5128, registerStateLIDs :5200-5257,
 *   registerStoreLIDs :5267-5290, Master::updateState :16180-16293,
 *   Master::loadDAEVectors :16303-16567, Master::loadDAEMatrices :16577-17268.
 * Option set as in or_bsim4_eval.c. */
#include <stdio.h>
#include "or_bsim4.h"

/* setupCapacitors_oldDAE, trnqsMod = 0, rgateMod != 3, rbodyMod = 0 */
void b4_setup_capacitors_old(b4inst* I, const xa_solver_state* ss) {
  const b4size* p = I->paramPtr;
  double ag0 = ss->pdt;
  if (!(ss->dcopFlag) && ss->initTranFlag && ss->newtonIter == 0) ag0 = 0.0;
  if (I->mode > 0) {
    I->qdrn -= I->qgdo;
    I->gcggb = (I->cggb + I->cgdo + I->cgso + p->cgbo) * ag0;
    I->gcgdb = (I->cgdb - I->cgdo) * ag0;
    I->gcgsb = (I->cgsb - I->cgso) * ag0;
    I->gcgbb = -(I->gcggb + I->gcgdb + I->gcgsb);
    I->gcdgb = (I->cdgb - I->cgdo) * ag0;
    I->gcsgb = -(I->cggb + I->cbgb + I->cdgb + I->cgso) * ag0;
    I->gcbgb = (I->cbgb - p->cgbo) * ag0;
    I->gcdgmb = I->gcsgmb = I->gcbgmb = 0.0;
    I->qgb = p->cgbo * I->vgb;
    I->qgate += I->qgdo + I->qgso + I->qgb;
    I->qbulk -= I->qgb;
    I->qsrc = -(I->qgate + I->qbulk + I->qdrn);
    I->gcddb = (I->cddb + I->capbd + I->cgdo) * ag0;
    I->gcdsb = I->cdsb * ag0;
    I->gcsdb = -(I->cgdb + I->cbdb + I->cddb) * ag0;
    I->gcssb = (I->capbs + I->cgso - (I->cgsb + I->cbsb + I->cdsb)) * ag0;
    I->gcdbb = -(I->gcdgb + I->gcddb + I->gcdsb + I->gcdgmb);
    I->gcsbb = -(I->gcsgb + I->gcsdb + I->gcssb + I->gcsgmb);
    I->gcbdb = (I->cbdb - I->capbd) * ag0;
    I->gcbsb = (I->cbsb - I->capbs) * ag0;
    I->gcdbdb = 0.0;
    I->gcsbsb = 0.0;
    I->gcbbb = -(I->gcbdb + I->gcbgb + I->gcbsb + I->gcbgmb);
    I->ggtg = I->ggtd = I->ggtb = I->ggts = 0.0;
    I->sxpart = 0.6;
    I->dxpart = 0.4;
    I->ddxpart_dVd = I->ddxpart_dVg = I->ddxpart_dVb = I->ddxpart_dVs = 0.0;
    I->dsxpart_dVd = I->dsxpart_dVg = I->dsxpart_dVb = I->dsxpart_dVs = 0.0;
  } else {
    I->qsrc = I->qdrn - I->qgso;
    I->gcggb = (I->cggb + I->cgdo + I->cgso + p->cgbo) * ag0;
    I->gcgdb = (I->cgsb - I->cgdo) * ag0;
    I->gcgsb = (I->cgdb - I->cgso) * ag0;
    I->gcgbb = -(I->gcggb + I->gcgdb + I->gcgsb);
    I->gcdgb = -(I->cggb + I->cbgb + I->cdgb + I->cgdo) * ag0;
    I->gcsgb = (I->cdgb - I->cgso) * ag0;
    I->gcbgb = (I->cbgb - p->cgbo) * ag0;
    I->gcdgmb = I->gcsgmb = I->gcbgmb = 0.0;
    I->qgb = p->cgbo * I->vgb;
    I->qgate += I->qgdo + I->qgso + I->qgb;
    I->qbulk -= I->qgb;
    I->qdrn = -(I->qgate + I->qbulk + I->qsrc);
    I->gcddb = (I->capbd + I->cgdo - (I->cgsb + I->cbsb + I->cdsb)) * ag0;
    I->gcdsb = -(I->cgdb + I->cbdb + I->cddb) * ag0;
    I->gcsdb = I->cdsb * ag0;
    I->gcssb = (I->cddb + I->capbs + I->cgso) * ag0;
    I->gcdbb = -(I->gcdgb + I->gcddb + I->gcdsb + I->gcdgmb);
    I->gcsbb = -(I->gcsgb + I->gcsdb + I->gcssb + I->gcsgmb);
    I->gcbdb = (I->cbsb - I->capbd) * ag0;
    I->gcbsb = (I->cbdb - I->capbs) * ag0;
    I->gcdbdb = 0.0;
    I->gcsbsb = 0.0;
    I->gcbbb = -(I->gcbgb + I->gcbdb + I->gcbsb + I->gcbgmb);
    I->ggtg = I->ggtd = I->ggtb = I->ggts = 0.0;
    I->sxpart = 0.4;
    I->dxpart = 0.6;
    I->ddxpart_dVd = I->ddxpart_dVg = I->ddxpart_dVb = I->ddxpart_dVs = 0.0;
    I->dsxpart_dVd = I->dsxpart_dVg = I->dsxpart_dVb = I->dsxpart_dVs = 0.0;
  }
}

/* setupCapacitors_newDAE, trnqsMod = 0, rgateMod != 3, rbodyMod = 0 */
void b4_setup_capacitors_new(b4inst* I) {
  const b4size* p = I->paramPtr;
  if (I->mode > 0) {
    I->CAPcggb = (I->cggb + I->cgdo + I->cgso + p->cgbo);
    I->CAPcgdb = (I->cgdb - I->cgdo);
    I->CAPcgsb = (I->cgsb - I->cgso);
    I->CAPcgbb = -(I->CAPcggb + I->CAPcgdb + I->CAPcgsb);
    I->CAPcdgb = (I->cdgb - I->cgdo);
    I->CAPcsgb = -(I->cggb + I->cbgb + I->cdgb + I->cgso);
    I->CAPcbgb = (I->cbgb - p->cgbo);
    I->CAPcdgmb = I->CAPcsgmb = I->CAPcbgmb = 0.0;
    I->CAPcddb = (I->cddb + I->capbd + I->cgdo);
    I->CAPcdsb = I->cdsb;
    I->CAPcsdb = -(I->cgdb + I->cbdb + I->cddb);
    I->CAPcssb = (I->capbs + I->cgso - (I->cgsb + I->cbsb + I->cdsb));
    I->CAPcdbb = -(I->CAPcdgb + I->CAPcddb + I->CAPcdsb + I->CAPcdgmb);
    I->CAPcsbb = -(I->CAPcsgb + I->CAPcsdb + I->CAPcssb + I->CAPcsgmb);
    I->CAPcbdb = (I->cbdb - I->capbd);
    I->CAPcbsb = (I->cbsb - I->capbs);
    I->CAPcdbdb = 0.0;
    I->CAPcsbsb = 0.0;
    I->CAPcbbb = -(I->CAPcbdb + I->CAPcbgb + I->CAPcbsb + I->CAPcbgmb);
  } else {
    I->CAPcggb = (I->cggb + I->cgdo + I->cgso + p->cgbo);
    I->CAPcgdb = (I->cgsb - I->cgdo);
    I->CAPcgsb = (I->cgdb - I->cgso);
    I->CAPcgbb = -(I->CAPcggb + I->CAPcgdb + I->CAPcgsb);
    I->CAPcdgb = -(I->cggb + I->cbgb + I->cdgb + I->cgdo);
    I->CAPcsgb = (I->cdgb - I->cgso);
    I->CAPcbgb = (I->cbgb - p->cgbo);
    I->CAPcdgmb = I->CAPcsgmb = I->CAPcbgmb = 0.0;
    I->CAPcddb = (I->capbd + I->cgdo - (I->cgsb + I->cbsb + I->cdsb));
    I->CAPcdsb = -(I->cgdb + I->cbdb + I->cddb);
    I->CAPcsdb = I->cdsb;
    I->CAPcssb = (I->cddb + I->capbs + I->cgso);
    I->CAPcdbb = -(I->CAPcdgb + I->CAPcddb + I->CAPcdsb + I->CAPcdgmb);
    I->CAPcsbb = -(I->CAPcsgb + I->CAPcsdb + I->CAPcssb + I->CAPcsgmb);
    I->CAPcbdb = (I->cbsb - I->capbd);
    I->CAPcbsb = (I->cbdb - I->capbs);
    I->CAPcdbdb = 0.0;
    I->CAPcsbsb = 0.0;
    I->CAPcbbb = -(I->CAPcbgb + I->CAPcbdb + I->CAPcbsb + I->CAPcbgmb);
  }
}

/* auxChargeCalculations (:12216-12347), trnqsMod = 0, rgateMod != 3 */
static void aux_charge(b4inst* I) {
  if (!I->ChargeComputationNeeded) {
    I->sxpart = (1.0 - (I->dxpart = (I->mode > 0) ? 0.4 : 0.6));
    I->ddxpart_dVd = I->ddxpart_dVg = I->ddxpart_dVb = I->ddxpart_dVs = 0.0;
    I->dsxpart_dVd = I->dsxpart_dVg = I->dsxpart_dVb = I->dsxpart_dVs = 0.0;
    I->Qeqqg_Jdxp = 0.0;
    I->Qeqqd_Jdxp = 0.0;
    I->Qeqqb_Jdxp = 0.0;
    I->Qeqqgmid_Jdxp = 0.0;
    I->Qeqqjs_Jdxp = 0.0;
    I->Qeqqjd_Jdxp = 0.0;
    I->gtau = 0.0;
  } else {
    I->Qeqqg_Jdxp = 0.0;
    if (!I->origFlag)
      I->Qeqqg_Jdxp = -I->CAPcggb * (I->vgb - I->vgb_orig) + I->CAPcgdb * (I->vbd - I->vbd_orig)
                      + I->CAPcgsb * (I->vbs - I->vbs_orig);
    I->Qeqqd_Jdxp = 0.0;
    if (!I->origFlag)
      I->Qeqqd_Jdxp = -I->CAPcdgb * (I->vgb - I->vgb_orig) - I->CAPcdgmb * (I->vgmb - I->vgmb_orig)
                      + (I->CAPcddb + I->CAPcdbdb) * (I->vbd - I->vbd_orig)
                      - I->CAPcdbdb * (I->vbd_jct - I->vbd_jct_orig) + I->CAPcdsb * (I->vbs - I->vbs_orig);
    I->Qeqqb_Jdxp = 0.0;
    if (!I->origFlag)
      I->Qeqqb_Jdxp = -I->CAPcbgb * (I->vgb - I->vgb_orig) - I->CAPcbgmb * (I->vgmb - I->vgmb_orig)
                      + I->CAPcbdb * (I->vbd - I->vbd_orig) + I->CAPcbsb * (I->vbs - I->vbs_orig);
    I->Qeqqgmid_Jdxp = 0.0;
  }
}

/* setupFVectorVars (:13525-13838), igcMod = igbMod = 0, rgateMod <= 1, rdsMod = 0 */
static void setup_fvector_vars(b4inst* I) {
  const b4model* m = I->model;
  const double dt = m->dtype;
  I->ceqdrn_Jdxp = 0.0; I->ceqbd_Jdxp = 0.0; I->ceqbs_Jdxp = 0.0;
  I->Istoteq_Jdxp = 0.0; I->Idtoteq_Jdxp = 0.0;
  I->Ibtoteq_Jdxp = 0.0; I->Igtoteq_Jdxp = 0.0;
  I->ceqgcrg_Jdxp = 0.0; I->ceqgstot_Jdxp = 0.0; I->ceqgdtot_Jdxp = 0.0;
  I->ceqjs_Jdxp = 0.0; I->ceqjd_Jdxp = 0.0;
  if (I->mode >= 0) {
    I->Gm = I->gm;
    I->Gmbs = I->gmbs;
    I->FwdSum = I->Gm + I->Gmbs;
    I->RevSum = 0.0;
    I->ceqdrn = dt * I->cdrain;
    I->ceqdrn_Jdxp = dt * (-I->gds * (I->vds - I->vds_orig) - I->Gm * (I->vgs - I->vgs_orig)
                           - I->Gmbs * (I->vbs - I->vbs_orig));
    I->ceqbd = dt * (I->csub + I->Igidl);
    I->ceqbd_Jdxp = dt * (-(I->gbds + I->ggidld) * (I->vds - I->vds_orig)
                          - (I->gbgs + I->ggidlg) * (I->vgs - I->vgs_orig)
                          - (I->gbbs + I->ggidlb) * (I->vbs - I->vbs_orig));
    I->ceqbs = dt * I->Igisl;
    I->ceqbs_Jdxp = dt * (+I->ggisls * (I->vds - I->vds_orig) - I->ggislg * (I->vgd - I->vgd_orig)
                          - I->ggislb * (I->vbd - I->vbd_orig));
    I->gbbdp = -(I->gbds);
    I->gbbsp = I->gbds + I->gbgs + I->gbbs;
    I->gbdpg = I->gbgs;
    I->gbdpdp = I->gbds;
    I->gbdpb = I->gbbs;
    I->gbdpsp = -(I->gbdpg + I->gbdpdp + I->gbdpb);
    I->gbspg = 0.0;
    I->gbspdp = 0.0;
    I->gbspb = 0.0;
    I->gbspsp = 0.0;
  } else {
    I->Gm = -I->gm;
    I->Gmbs = -I->gmbs;
    I->FwdSum = 0.0;
    I->RevSum = -(I->Gm + I->Gmbs);
    I->ceqdrn = -dt * I->cdrain;
    I->ceqdrn_Jdxp = -dt * (+I->gds * (I->vds - I->vds_orig) + I->Gm * (I->vgd - I->vgd_orig)
                            + I->Gmbs * (I->vbd - I->vbd_orig));
    I->ceqbs = dt * (I->csub + I->Igisl);
    I->ceqbs_Jdxp = dt * (+(I->gbds + I->ggisls) * (I->vds - I->vds_orig)
                          - (I->gbgs + I->ggislg) * (I->vgd - I->vgd_orig)
                          - (I->gbbs + I->ggislb) * (I->vbd - I->vbd_orig));
    I->ceqbd = dt * I->Igidl;
    I->ceqbd_Jdxp = dt * (-I->ggidld * (I->vds - I->vds_orig) - I->ggidlg * (I->vgs - I->vgs_orig)
                          - I->ggidlb * (I->vbs - I->vbs_orig));
    I->gbbsp = -(I->gbds);
    I->gbbdp = I->gbds + I->gbgs + I->gbbs;
    I->gbdpg = 0.0;
    I->gbdpsp = 0.0;
    I->gbdpb = 0.0;
    I->gbdpdp = 0.0;
    I->gbspg = I->gbgs;
    I->gbspsp = I->gbds;
    I->gbspb = I->gbbs;
    I->gbspdp = -(I->gbspg + I->gbspsp + I->gbspb);
  }
  I->gIstotg = I->gIstotd = I->gIstots = I->gIstotb = I->Istoteq = 0.0;
  I->gIdtotg = I->gIdtotd = I->gIdtots = I->gIdtotb = I->Idtoteq = 0.0;
  I->gIbtotg = I->gIbtotd = I->gIbtots = I->gIbtotb = I->Ibtoteq = 0.0;
  I->gIgtotg = I->gIgtotd = I->gIgtots = I->gIgtotb = I->Igtoteq = 0.0;
  I->ceqgcrg = I->gcrg = I->gcrgd = I->gcrgg = I->gcrgs = I->gcrgb = 0.0;
  I->gstot = I->gstotd = I->gstotg = I->gstots = I->gstotb = I->ceqgstot = 0.0;
  I->gdtot = I->gdtotd = I->gdtotg = I->gdtots = I->gdtotb = I->ceqgdtot = 0.0;
  if (m->dtype > 0) {
    I->ceqjs = (I->cbs);
    I->ceqjs_Jdxp = (-I->gbs * (I->vbs_jct - I->vbs_jct_orig));
    I->ceqjd = (I->cbd);
    I->ceqjd_Jdxp = (-I->gbd * (I->vbd_jct - I->vbd_jct_orig));
  } else {
    I->ceqjs = -(I->cbs);
    I->ceqjs_Jdxp = (I->gbs * (I->vbs_jct - I->vbs_jct_orig));
    I->ceqjd = -(I->cbd);
    I->ceqjd_Jdxp = (I->gbd * (I->vbd_jct - I->vbd_jct_orig));
    I->ceqgcrg = -I->ceqgcrg;
    I->ceqgcrg_Jdxp = -I->ceqgcrg_Jdxp;
  }
}

/* ---------------------------------------------------------------------- */
/* parameter table: B4M rows, the TYPE pseudo-parameter, then B4I rows */
enum {
#define B4M(NAME, mem, def) B4S_##mem,
#include "../include/xa_bsim4_params.def"
#undef B4M
  B4S_TYPE, B4S_NM
};
enum {
#define B4I(NAME, mem, def) B4IS_##mem,
#include "../include/xa_bsim4_params.def"
#undef B4I
  B4IS_NI
};
static const or_param mpar[] = {
#define B4M(NAME, mem, def) {#NAME, B4S_##mem, (double)(def), 0},
#include "../include/xa_bsim4_params.def"
#undef B4M
  {"TYPE", B4S_TYPE, 1.0, 0},
};
static const or_param ipar[] = {
#define B4I(NAME, mem, def) {#NAME, B4IS_##mem, (double)(def), 0},
#include "../include/xa_bsim4_params.def"
#undef B4I
};

static int is_temp_name(const char* s) { return !strcmp(s, "TNOM") || !strcmp(s, "TEMP"); }
static or_param mpar_t[sizeof(mpar) / sizeof(mpar[0])];
static or_param ipar_t[sizeof(ipar) / sizeof(ipar[0])];

static int setup(struct or_group_s* g) {
  int k, rc;
  int64_t i;
  b4model* mods = (b4model*)calloc(g->nmodels, sizeof(b4model));
  b4inst* in = (b4inst*)calloc(g->n > 0 ? g->n : 1, sizeof(b4inst));
  g->mdata = mods;
  g->inst = in;
  for (k = 0; k < g->nmodels; ++k) {
    if (g->models[k].level != 54 && g->models[k].level != 14)
      return or_fail(XA_ERR_UNSUPPORTED, "bsim4: level must be 14 or 54");
    if ((rc = b4_model_init(&mods[k], &g->models[k], &g->opt)) != XA_OK) return rc;
  }
  for (i = 0; i < g->n; ++i) {
    in[i].model = &mods[g->inst_model[i]];
    if ((rc = b4_instance_init(g, &in[i], i)) != XA_OK) return rc;
  }
  return XA_OK;
}

static void destroy(struct or_group_s* g) {
  int k, s;
  b4model* mods = (b4model*)g->mdata;
  if (!mods) return;
  for (k = 0; k < g->nmodels; ++k) {
    for (s = 0; s < mods[k].nsizes; ++s) free(mods[k].sizes[s]);
    free(mods[k].sizes);
  }
}

static void register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes, const int32_t* sta,
                          const int32_t* sto) {
  b4inst* I = &((b4inst*)g->inst)[i];
  int loc = 4;
  I->li_Drain = nodes[0];
  I->li_GateExt = nodes[1];
  I->li_Source = nodes[2];
  I->li_Body = nodes[3];
  I->li_DrainPrime = I->drainMOSFET_B4Exists ? nodes[loc++] : I->li_Drain;
  I->li_SourcePrime = I->sourceMOSFET_B4Exists ? nodes[loc++] : I->li_Source;
  I->li_GatePrime = I->li_GateExt;
  I->li_GateMid = I->li_GateExt;
  I->li_BodyPrime = I->li_Body;
  I->li_SourceBody = I->li_Body;
  I->li_DrainBody = I->li_Body;
  I->li_state_qb = sta ? sta[0] : -1;
  I->li_state_qg = sta ? sta[1] : -1;
  I->li_state_qd = sta ? sta[2] : -1;
  I->li_store_vbd = sto[0];
  I->li_store_vbs = sto[1];
  I->li_store_vgs = sto[2];
  I->li_store_vds = sto[3];
  I->li_store_vges = sto[4];
  I->li_store_vgms = sto[5];
  I->li_store_vdes = sto[6];
  I->li_store_vses = sto[7];
  I->li_store_vdbs = sto[8];
  I->li_store_vsbs = sto[9];
  I->li_store_vdbd = sto[10];
  I->li_store_vged = sto[11];
  I->li_store_vgmd = sto[12];
  I->li_store_von = sto[13];
}

/* (row, col) node pairs stamped by loadDAEMatrices for this option set */
static void pair_list(const b4inst* I, int32_t pr[22], int32_t pc[22]) {
  const int D = I->li_Drain, S = I->li_Source, DP = I->li_DrainPrime, SP = I->li_SourcePrime;
  const int GP = I->li_GatePrime, BP = I->li_BodyPrime;
  const int r[22] = {GP, GP, GP, GP, DP, DP, DP, DP, DP, D, D, SP, SP, SP, SP, SP, S, S, BP, BP, BP, BP};
  const int c[22] = {GP, DP, SP, BP, DP, D, GP, SP, BP, DP, D, DP, GP, SP, S, BP, SP, S, DP, GP, SP, BP};
  int k;
  for (k = 0; k < 22; ++k) { pr[k] = r[k]; pc[k] = c[k]; }
}

static int64_t jac_pairs(struct or_group_s* g, int64_t i, int32_t* r, int32_t* c) {
  const b4inst* I = &((b4inst*)g->inst)[i];
  int32_t pr[22], pc[22];
  int64_t k, n = 0;
  pair_list(I, pr, pc);
  for (k = 0; k < 22; ++k) {
    if (pr[k] < 0 || pc[k] < 0) continue;
    if (r) { r[n] = pr[k]; c[n] = pc[k]; }
    ++n;
  }
  return n;
}

static int register_jac(struct or_group_s* g, int64_t i) {
  b4inst* I = &((b4inst*)g->inst)[i];
  const int D = I->li_Drain, S = I->li_Source, DP = I->li_DrainPrime, SP = I->li_SourcePrime;
  const int GP = I->li_GatePrime, BP = I->li_BodyPrime;
  I->s_GPgp = or_slot(g, GP, GP); I->s_GPdp = or_slot(g, GP, DP);
  I->s_GPsp = or_slot(g, GP, SP); I->s_GPbp = or_slot(g, GP, BP);
  I->s_DPdp = or_slot(g, DP, DP); I->s_DPd = or_slot(g, DP, D);
  I->s_DPgp = or_slot(g, DP, GP); I->s_DPsp = or_slot(g, DP, SP);
  I->s_DPbp = or_slot(g, DP, BP); I->s_Ddp = or_slot(g, D, DP); I->s_Dd = or_slot(g, D, D);
  I->s_SPdp = or_slot(g, SP, DP); I->s_SPgp = or_slot(g, SP, GP);
  I->s_SPsp = or_slot(g, SP, SP); I->s_SPs = or_slot(g, SP, S);
  I->s_SPbp = or_slot(g, SP, BP); I->s_Ssp = or_slot(g, S, SP); I->s_Ss = or_slot(g, S, S);
  I->s_BPdp = or_slot(g, BP, DP); I->s_BPgp = or_slot(g, BP, GP);
  I->s_BPsp = or_slot(g, BP, SP); I->s_BPbp = or_slot(g, BP, BP);
  {
    const int64_t* s = &I->s_GPgp;
    int k;
    for (k = 0; k < 22; ++k) if (s[k] < -1) return XA_ERR_ARG;
  }
  return XA_OK;
}

/* Master::updateState (:16180-16293) */
static int update_state(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  b4inst* in = (b4inst*)g->inst;
  int ok = 1;
  /* instances are independent here (the reference's _OMP loop,
   * N_DEV_MOSFET_B4.C:16184); one thread is its serial order */
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1) reduction(&& : ok)
  for (i = 0; i < g->n; ++i) {
    b4inst* I = &in[i];
    double* sto = v->nextSto;
    double* sta = v->nextSta;
    ok = b4_update_intermediate(I, ss, &g->opt, v) && ok;
    sto[I->li_store_vbd] = I->vbd;
    sto[I->li_store_vbs] = I->vbs;
    sto[I->li_store_vgs] = I->vgs;
    sto[I->li_store_vds] = I->vds;
    sto[I->li_store_vges] = I->vges;
    sto[I->li_store_vgms] = I->vgms;
    sto[I->li_store_vdes] = I->vdes;
    sto[I->li_store_vses] = I->vses;
    sto[I->li_store_vdbs] = I->vdbs;
    sto[I->li_store_vsbs] = I->vsbs;
    sto[I->li_store_vdbd] = I->vdbd;
    sto[I->li_store_von] = I->von;
    I->qg = I->qgate;
    I->qd = I->qdrn - I->qbd;
    I->qb = I->qbulk + I->qbd + I->qbs;
    if (sta) {
      if (I->li_state_qg >= 0) sta[I->li_state_qg] = I->qg;
      if (I->li_state_qd >= 0) sta[I->li_state_qd] = I->qd;
      if (I->li_state_qb >= 0) sta[I->li_state_qb] = I->qb;
    }
    if (!(ss->dcopFlag) && ss->initTranFlag && ss->newtonIter == 0 && v->currSta) {
      double* cs = v->currSta;
      if (I->li_state_qg >= 0) cs[I->li_state_qg] = I->qgate;
      if (I->li_state_qd >= 0) cs[I->li_state_qd] = I->qdrn - I->qbd;
      if (I->li_state_qb >= 0) cs[I->li_state_qb] = I->qbulk + I->qbd + I->qbs;
    }
  }
  return ok ? XA_OK : XA_ERR_DEVICE;
}

/* Master::loadDAEVectors (:16303-16567) */
static int load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  b4inst* in = (b4inst*)g->inst;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    b4inst* mi = &in[i];
    const double Mp = mi->numberParallel;
    double Qeqqg, Qeqqb, Qeqqd, Qeqqgmid = 0.0;
    setup_fvector_vars(mi);
    if (v->f) {
      OR_ADD(v->f, mi->li_DrainPrime, -(mi->ceqjd - mi->ceqbd - mi->ceqdrn + mi->Idtoteq) * Mp);
      OR_SUB(v->f, mi->li_GatePrime, -(-mi->ceqgcrg + mi->Igtoteq) * Mp);
      OR_ADD(v->f, mi->li_BodyPrime, -(mi->ceqbd + mi->ceqbs - mi->ceqjd - mi->ceqjs + mi->Ibtoteq) * Mp);
      OR_ADD(v->f, mi->li_SourcePrime, -(mi->ceqdrn - mi->ceqbs + mi->ceqjs + mi->Istoteq) * Mp);
      if (mi->drainMOSFET_B4Exists) {
        OR_ADD(v->f, mi->li_Drain, -(-mi->Idrain) * Mp);
        OR_ADD(v->f, mi->li_DrainPrime, -(mi->Idrain) * Mp);
      }
      if (mi->sourceMOSFET_B4Exists) {
        OR_ADD(v->f, mi->li_Source, -(-mi->Isource) * Mp);
        OR_ADD(v->f, mi->li_SourcePrime, -(+mi->Isource) * Mp);
      }
    }
    if (g->opt.voltageLimiterFlag && !mi->origFlag && v->dFdxdVp) {
      OR_ADD(v->dFdxdVp, mi->li_DrainPrime,
             (mi->ceqjd_Jdxp - mi->ceqbd_Jdxp - mi->ceqdrn_Jdxp + mi->Idtoteq_Jdxp) * Mp);
      OR_SUB(v->dFdxdVp, mi->li_GatePrime, (-mi->ceqgcrg_Jdxp + mi->Igtoteq_Jdxp) * Mp);
      OR_ADD(v->dFdxdVp, mi->li_BodyPrime,
             (mi->ceqbd_Jdxp + mi->ceqbs_Jdxp - mi->ceqjd_Jdxp - mi->ceqjs_Jdxp + mi->Ibtoteq_Jdxp) * Mp);
      OR_ADD(v->dFdxdVp, mi->li_SourcePrime,
             (mi->ceqdrn_Jdxp - mi->ceqbs_Jdxp + mi->ceqjs_Jdxp + mi->Istoteq_Jdxp) * Mp);
    }
    aux_charge(mi);
    if (mi->model->dtype > 0) {
      Qeqqg = mi->qg;
      Qeqqd = mi->qd;
      Qeqqb = mi->qb;
    } else {
      Qeqqg = -mi->qg;
      Qeqqd = -mi->qd;
      Qeqqb = -mi->qb;
    }
    if (v->q) {
      OR_ADD(v->q, mi->li_DrainPrime, -(-Qeqqd) * Mp);
      OR_SUB(v->q, mi->li_GatePrime, -(Qeqqg) * Mp);
      OR_ADD(v->q, mi->li_BodyPrime, -(-Qeqqb) * Mp);
      OR_ADD(v->q, mi->li_SourcePrime, -(+Qeqqg + Qeqqb + Qeqqd + Qeqqgmid) * Mp);
    }
    if (g->opt.voltageLimiterFlag && !mi->origFlag && v->dQdxdVp) {
      OR_ADD(v->dQdxdVp, mi->li_DrainPrime, (-mi->Qeqqd_Jdxp) * Mp);
      OR_SUB(v->dQdxdVp, mi->li_GatePrime, (mi->Qeqqg_Jdxp) * Mp);
      OR_ADD(v->dQdxdVp, mi->li_BodyPrime, (-mi->Qeqqb_Jdxp) * Mp);
      OR_ADD(v->dQdxdVp, mi->li_SourcePrime,
             (+mi->Qeqqg_Jdxp + mi->Qeqqb_Jdxp + mi->Qeqqd_Jdxp + mi->Qeqqgmid_Jdxp) * Mp);
    }
  }
  return XA_OK;
}

/* Master::loadDAEMatrices (:16577-17268) */
static int load_matrices(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  b4inst* in = (b4inst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    b4inst* mi = &in[i];
    const double Mp = mi->numberParallel;
    double T1;
    double* F = v->dFdx;
    double* Q = v->dQdx;
    mi->gjbd = mi->gbd;
    mi->gjbs = mi->gbs;
    mi->gdpr = mi->drainConductance;
    mi->gspr = mi->sourceConductance;
    mi->geltd = mi->grgeltd;
    T1 = mi->qdef * mi->gtau;
    if (F) {
      OR_ADD(F, mi->s_GPgp, (-mi->ggtg + mi->gIgtotg) * Mp);
      OR_ADD(F, mi->s_GPdp, (-mi->ggtd + mi->gIgtotd) * Mp);
      OR_ADD(F, mi->s_GPsp, (-mi->ggts + mi->gIgtots) * Mp);
      OR_ADD(F, mi->s_GPbp, (-mi->ggtb + mi->gIgtotb) * Mp);
      OR_ADD(F, mi->s_DPdp, (mi->gdpr + mi->gds + mi->gbd + T1 * mi->ddxpart_dVd - mi->gdtotd + mi->RevSum
                             + mi->gbdpdp + mi->dxpart * mi->ggtd - mi->gIdtotd) * Mp);
      OR_SUB(F, mi->s_DPd, (mi->gdpr + mi->gdtot) * Mp);
      OR_ADD(F, mi->s_DPgp, (mi->Gm - mi->gdtotg + mi->gbdpg - mi->gIdtotg + mi->dxpart * mi->ggtg
                             + T1 * mi->ddxpart_dVg) * Mp);
      OR_SUB(F, mi->s_DPsp, (mi->gds + mi->gdtots - mi->dxpart * mi->ggts + mi->gIdtots - T1 * mi->ddxpart_dVs
                             + mi->FwdSum - mi->gbdpsp) * Mp);
      OR_SUB(F, mi->s_DPbp, (mi->gjbd + mi->gdtotb - mi->Gmbs - mi->gbdpb + mi->gIdtotb - T1 * mi->ddxpart_dVb
                             - mi->dxpart * mi->ggtb) * Mp);
      OR_SUB(F, mi->s_Ddp, (mi->gdpr - mi->gdtotd) * Mp);
      OR_ADD(F, mi->s_Dd, (mi->gdpr + mi->gdtot) * Mp);
      OR_SUB(F, mi->s_SPdp, (mi->gds + mi->gstotd + mi->RevSum - mi->gbspdp - T1 * mi->dsxpart_dVd
                             - mi->sxpart * mi->ggtd + mi->gIstotd) * Mp);
      OR_ADD(F, mi->s_SPgp, (-mi->Gm - mi->gstotg + mi->gbspg + mi->sxpart * mi->ggtg + T1 * mi->dsxpart_dVg
                             - mi->gIstotg) * Mp);
      OR_ADD(F, mi->s_SPsp, (mi->gspr + mi->gds + mi->gbs + T1 * mi->dsxpart_dVs - mi->gstots + mi->FwdSum
                             + mi->gbspsp + mi->sxpart * mi->ggts - mi->gIstots) * Mp);
      OR_SUB(F, mi->s_SPs, (mi->gspr + mi->gstot) * Mp);
      OR_SUB(F, mi->s_SPbp, (mi->gjbs + mi->gstotb + mi->Gmbs - mi->gbspb - mi->sxpart * mi->ggtb
                             - T1 * mi->dsxpart_dVb + mi->gIstotb) * Mp);
      OR_SUB(F, mi->s_Ssp, (mi->gspr - mi->gstots) * Mp);
      OR_ADD(F, mi->s_Ss, (mi->gspr + mi->gstot) * Mp);
      OR_ADD(F, mi->s_BPdp, (-mi->gjbd + mi->gbbdp - mi->gIbtotd) * Mp);
      OR_ADD(F, mi->s_BPgp, (-mi->gbgs - mi->gIbtotg) * Mp);
      OR_ADD(F, mi->s_BPsp, (-mi->gjbs + mi->gbbsp - mi->gIbtots) * Mp);
      OR_ADD(F, mi->s_BPbp, (mi->gjbd + mi->gjbs - mi->gbbs - mi->gIbtotb) * Mp);
      OR_ADD(F, mi->s_DPdp, (mi->ggidld) * Mp);
      OR_ADD(F, mi->s_DPgp, (mi->ggidlg) * Mp);
      OR_SUB(F, mi->s_DPsp, ((mi->ggidlg + mi->ggidld + mi->ggidlb)) * Mp);
      OR_ADD(F, mi->s_DPbp, (mi->ggidlb) * Mp);
      OR_SUB(F, mi->s_BPdp, (mi->ggidld) * Mp);
      OR_SUB(F, mi->s_BPgp, (mi->ggidlg) * Mp);
      OR_ADD(F, mi->s_BPsp, ((mi->ggidlg + mi->ggidld + mi->ggidlb)) * Mp);
      OR_SUB(F, mi->s_BPbp, (mi->ggidlb) * Mp);
      OR_SUB(F, mi->s_SPdp, ((mi->ggisls + mi->ggislg + mi->ggislb)) * Mp);
      OR_ADD(F, mi->s_SPgp, (mi->ggislg) * Mp);
      OR_ADD(F, mi->s_SPsp, (mi->ggisls) * Mp);
      OR_ADD(F, mi->s_SPbp, (mi->ggislb) * Mp);
      OR_ADD(F, mi->s_BPdp, ((mi->ggislg + mi->ggisls + mi->ggislb)) * Mp);
      OR_SUB(F, mi->s_BPgp, (mi->ggislg) * Mp);
      OR_SUB(F, mi->s_BPsp, (mi->ggisls) * Mp);
      OR_SUB(F, mi->s_BPbp, (mi->ggislb) * Mp);
    }
    if (!(ss->dcopFlag) && ss->initTranFlag && ss->newtonIter == 0) {
      /* q is zero on the first transient iteration */
    } else if (Q) {
      OR_ADD(Q, mi->s_GPgp, (mi->CAPcggb) * Mp);
      OR_ADD(Q, mi->s_GPdp, (mi->CAPcgdb) * Mp);
      OR_ADD(Q, mi->s_GPsp, (mi->CAPcgsb) * Mp);
      OR_ADD(Q, mi->s_GPbp, (mi->CAPcgbb) * Mp);
      OR_ADD(Q, mi->s_DPdp, (mi->CAPcddb) * Mp);
      OR_ADD(Q, mi->s_DPgp, (+mi->CAPcdgb) * Mp);
      OR_SUB(Q, mi->s_DPsp, (-mi->CAPcdsb) * Mp);
      OR_SUB(Q, mi->s_DPbp, (-mi->CAPcdbb) * Mp);
      OR_SUB(Q, mi->s_SPdp, (-mi->CAPcsdb) * Mp);
      OR_ADD(Q, mi->s_SPgp, (mi->CAPcsgb) * Mp);
      OR_ADD(Q, mi->s_SPsp, (mi->CAPcssb) * Mp);
      OR_SUB(Q, mi->s_SPbp, (-mi->CAPcsbb) * Mp);
      OR_ADD(Q, mi->s_BPdp, (mi->CAPcbdb) * Mp);
      OR_ADD(Q, mi->s_BPgp, (mi->CAPcbgb) * Mp);
      OR_ADD(Q, mi->s_BPsp, (mi->CAPcbsb) * Mp);
      OR_ADD(Q, mi->s_BPbp, (mi->CAPcbbb) * Mp);
    }
  }
  return XA_OK;
}

static int read_var(struct or_group_s* g, const char* name, double* out) {
  int64_t i;
  b4inst* in = (b4inst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    const b4inst* I = &in[i];
    const b4size* p = I->paramPtr;
    if (!strcmp(name, "cdrain")) out[i] = I->cdrain;
    else if (!strcmp(name, "gm")) out[i] = I->gm;
    else if (!strcmp(name, "gds")) out[i] = I->gds;
    else if (!strcmp(name, "gmbs")) out[i] = I->gmbs;
    else if (!strcmp(name, "von")) out[i] = I->von;
    else if (!strcmp(name, "vdsat")) out[i] = I->vdsat;
    else if (!strcmp(name, "qgate")) out[i] = I->qg;
    else if (!strcmp(name, "qb")) out[i] = I->qb;
    else if (!strcmp(name, "qd")) out[i] = I->qd;
    else if (!strcmp(name, "mode")) out[i] = I->mode;
    else if (!strcmp(name, "origFlag")) out[i] = I->origFlag;
    else if (!strcmp(name, "vth0")) out[i] = I->vth0;
    else if (!strcmp(name, "leff")) out[i] = p->leff;
    else if (!strcmp(name, "weff")) out[i] = p->weff;
    else if (!strcmp(name, "u0temp")) out[i] = I->u0temp;
    else if (!strcmp(name, "vgs")) out[i] = I->vgs;
    else if (!strcmp(name, "vds")) out[i] = I->vds;
    else if (!strcmp(name, "vbs")) out[i] = I->vbs;
    else if (!strcmp(name, "Vgsteff")) out[i] = I->dcVgsteff;
    else if (!strcmp(name, "Abulk")) out[i] = I->dcAbulk;
    else if (!strcmp(name, "Vdseff")) out[i] = I->dcVdseff;
    else if (!strcmp(name, "ueff")) out[i] = I->ueff;
    else if (!strcmp(name, "EsatL")) out[i] = I->EsatL;
    else if (!strcmp(name, "vsattemp")) out[i] = I->vsattemp;
    else if (!strcmp(name, "vtm")) out[i] = I->model->vtm;
    else if (!strcmp(name, "coxe")) out[i] = I->model->coxe;
    else if (!strcmp(name, "qsrc")) out[i] = I->qsrc;
    else if (!strcmp(name, "qdrn")) out[i] = I->qdrn;
    else if (!strcmp(name, "Igidl")) out[i] = I->Igidl;
    else if (!strcmp(name, "weffCJ")) out[i] = p->weffCJ;
    else return or_fail(XA_ERR_ARG, "bsim4: unknown intermediate");
  }
  return XA_OK;
}

/* The generic plumbing matches names exactly after upper-casing both; the
 * table holds the reference's spellings, temperature flags filled here. */
static int tables_ready = 0;
static void prepare_tables(void) {
  size_t k;
  if (tables_ready) return;
  for (k = 0; k < sizeof(mpar) / sizeof(mpar[0]); ++k) {
    mpar_t[k] = mpar[k];
    mpar_t[k].is_temp = is_temp_name(mpar[k].name);
  }
  for (k = 0; k < sizeof(ipar) / sizeof(ipar[0]); ++k) {
    ipar_t[k] = ipar[k];
    ipar_t[k].is_temp = is_temp_name(ipar[k].name);
  }
  tables_ready = 1;
}

static int setup_entry(struct or_group_s* g) {
  prepare_tables();
  return setup(g);
}

const or_model_ops or_bsim4_ops = {
  "bsim4", mpar_t, sizeof(mpar) / sizeof(mpar[0]), B4S_NM, ipar_t,
  sizeof(ipar) / sizeof(ipar[0]), B4IS_NI, setup_entry, register_lids, jac_pairs, register_jac,
  update_state, NULL, load_vectors, load_matrices, read_var, destroy, sizeof(b4inst),
};

/* tables must be ready before the first add_model, which precedes setup */
__attribute__((constructor)) static void b4_tables_init(void) { prepare_tables(); }
