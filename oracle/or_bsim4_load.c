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

/* setupCapacitors_oldDAE (:12616-13031): the charges the state vector
 * takes, the partition of the channel charge and the NQS conductances (the
 * ag0-scaled capacitances are the old DAE's and kept for completeness) */
void b4_setup_capacitors_old(b4inst* I, const xa_solver_state* ss) {
  const b4size* p = I->paramPtr;
  const b4model* m = I->model;
  double ag0 = ss->pdt, T0, Cdd, Csd, Cdg, Csg, Cds, Css;
  if (!(ss->dcopFlag) && ss->initTranFlag && ss->newtonIter == 0) ag0 = 0.0;
  if (I->mode > 0) {
    if (I->itrnqsMod == 0) {
      I->qdrn -= I->qgdo;
      if (I->irgateMod == 3) {
        I->gcgmgmb = (I->cgdo + I->cgso + p->cgbo) * ag0;
        I->gcgmdb = -I->cgdo * ag0;
        I->gcgmsb = -I->cgso * ag0;
        I->gcgmbb = -p->cgbo * ag0;
        I->gcdgmb = I->gcgmdb;
        I->gcsgmb = I->gcgmsb;
        I->gcbgmb = I->gcgmbb;
        I->gcggb = I->cggb * ag0;
        I->gcgdb = I->cgdb * ag0;
        I->gcgsb = I->cgsb * ag0;
        I->gcgbb = -(I->gcggb + I->gcgdb + I->gcgsb);
        I->gcdgb = I->cdgb * ag0;
        I->gcsgb = -(I->cggb + I->cbgb + I->cdgb) * ag0;
        I->gcbgb = I->cbgb * ag0;
        I->qgmb = p->cgbo * I->vgmb;
        I->qgmid = I->qgdo + I->qgso + I->qgmb;
        I->qbulk -= I->qgmb;
        I->qsrc = -(I->qgate + I->qgmid + I->qbulk + I->qdrn);
      } else {
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
      }
      I->gcddb = (I->cddb + I->capbd + I->cgdo) * ag0;
      I->gcdsb = I->cdsb * ag0;
      I->gcsdb = -(I->cgdb + I->cbdb + I->cddb) * ag0;
      I->gcssb = (I->capbs + I->cgso - (I->cgsb + I->cbsb + I->cdsb)) * ag0;
      if (!I->irbodyMod) {
        I->gcdbb = -(I->gcdgb + I->gcddb + I->gcdsb + I->gcdgmb);
        I->gcsbb = -(I->gcsgb + I->gcsdb + I->gcssb + I->gcsgmb);
        I->gcbdb = (I->cbdb - I->capbd) * ag0;
        I->gcbsb = (I->cbsb - I->capbs) * ag0;
        I->gcdbdb = 0.0;
        I->gcsbsb = 0.0;
      } else {
        I->gcdbb = -(I->cddb + I->cdgb + I->cdsb) * ag0;
        I->gcsbb = -(I->gcsgb + I->gcsdb + I->gcssb + I->gcsgmb) + I->capbs * ag0;
        I->gcbdb = I->cbdb * ag0;
        I->gcbsb = I->cbsb * ag0;
        I->gcdbdb = -I->capbd * ag0;
        I->gcsbsb = -I->capbs * ag0;
      }
      I->gcbbb = -(I->gcbdb + I->gcbgb + I->gcbsb + I->gcbgmb);
      I->ggtg = I->ggtd = I->ggtb = I->ggts = 0.0;
      I->sxpart = 0.6;
      I->dxpart = 0.4;
      I->ddxpart_dVd = I->ddxpart_dVg = I->ddxpart_dVb = I->ddxpart_dVs = 0.0;
      I->dsxpart_dVd = I->dsxpart_dVg = I->dsxpart_dVb = I->dsxpart_dVs = 0.0;
    } else {
      I->qcheq = I->qchqs;
      I->CoxWL = m->coxe * p->weffCV * I->nf * p->leffCV;
      T0 = I->qdef * I->ScalingFactor / I->CoxWL;
      I->ggtg = I->gtg = T0 * I->gcrgg;
      I->ggtd = I->gtd = T0 * I->gcrgd;
      I->ggts = I->gts = T0 * I->gcrgs;
      I->ggtb = I->gtb = T0 * I->gcrgb;
      I->gqdef = I->ScalingFactor * ag0;
      I->gcqgb = I->cqgb * ag0;
      I->gcqdb = I->cqdb * ag0;
      I->gcqsb = I->cqsb * ag0;
      I->gcqbb = I->cqbb * ag0;
      if (fabs(I->qcheq) <= 1.0e-5 * I->CoxWL) {
        if (M(xpart) < 0.5) I->dxpart = 0.4;
        else if (M(xpart) > 0.5) I->dxpart = 0.0;
        else I->dxpart = 0.5;
        I->ddxpart_dVd = I->ddxpart_dVg = I->ddxpart_dVb = I->ddxpart_dVs = 0.0;
      } else {
        I->dxpart = I->qdrn / I->qcheq;
        Cdd = I->cddb;
        Csd = -(I->cgdb + I->cddb + I->cbdb);
        I->ddxpart_dVd = (Cdd - I->dxpart * (Cdd + Csd)) / I->qcheq;
        Cdg = I->cdgb;
        Csg = -(I->cggb + I->cdgb + I->cbgb);
        I->ddxpart_dVg = (Cdg - I->dxpart * (Cdg + Csg)) / I->qcheq;
        Cds = I->cdsb;
        Css = -(I->cgsb + I->cdsb + I->cbsb);
        I->ddxpart_dVs = (Cds - I->dxpart * (Cds + Css)) / I->qcheq;
        I->ddxpart_dVb = -(I->ddxpart_dVd + I->ddxpart_dVg + I->ddxpart_dVs);
      }
      I->sxpart = 1.0 - I->dxpart;
      I->dsxpart_dVd = -I->ddxpart_dVd;
      I->dsxpart_dVg = -I->ddxpart_dVg;
      I->dsxpart_dVs = -I->ddxpart_dVs;
      I->dsxpart_dVb = -(I->dsxpart_dVd + I->dsxpart_dVg + I->dsxpart_dVs);
      if (I->irgateMod == 3) {
        I->gcgmgmb = (I->cgdo + I->cgso + p->cgbo) * ag0;
        I->gcgmdb = -I->cgdo * ag0;
        I->gcgmsb = -I->cgso * ag0;
        I->gcgmbb = -p->cgbo * ag0;
        I->gcdgmb = I->gcgmdb;
        I->gcsgmb = I->gcgmsb;
        I->gcbgmb = I->gcgmbb;
        I->gcdgb = I->gcsgb = I->gcbgb = 0.0;
        I->gcggb = I->gcgdb = I->gcgsb = I->gcgbb = 0.0;
        I->qgmb = p->cgbo * I->vgmb;
        I->qgmid = I->qgdo + I->qgso + I->qgmb;
        I->qgate = 0.0;
        I->qbulk = -I->qgmb;
        I->qdrn = -I->qgdo;
        I->qsrc = -(I->qgmid + I->qbulk + I->qdrn);
      } else {
        I->gcggb = (I->cgdo + I->cgso + p->cgbo) * ag0;
        I->gcgdb = -I->cgdo * ag0;
        I->gcgsb = -I->cgso * ag0;
        I->gcgbb = -p->cgbo * ag0;
        I->gcdgb = I->gcgdb;
        I->gcsgb = I->gcgsb;
        I->gcbgb = I->gcgbb;
        I->gcdgmb = I->gcsgmb = I->gcbgmb = 0.0;
        I->qgb = p->cgbo * I->vgb;
        I->qgate = I->qgdo + I->qgso + I->qgb;
        I->qbulk = -I->qgb;
        I->qdrn = -I->qgdo;
        I->qsrc = -(I->qgate + I->qbulk + I->qdrn);
      }
      I->gcddb = (I->capbd + I->cgdo) * ag0;
      I->gcdsb = I->gcsdb = 0.0;
      I->gcssb = (I->capbs + I->cgso) * ag0;
      if (!I->irbodyMod) {
        I->gcdbb = -(I->gcdgb + I->gcddb + I->gcdgmb);
        I->gcsbb = -(I->gcsgb + I->gcssb + I->gcsgmb);
        I->gcbdb = -I->capbd * ag0;
        I->gcbsb = -I->capbs * ag0;
        I->gcdbdb = 0.0;
        I->gcsbsb = 0.0;
      } else {
        I->gcdbb = I->gcsbb = I->gcbdb = I->gcbsb = 0.0;
        I->gcdbdb = -I->capbd * ag0;
        I->gcsbsb = -I->capbs * ag0;
      }
      I->gcbbb = -(I->gcbdb + I->gcbgb + I->gcbsb + I->gcbgmb);
    }
  } else {
    if (I->itrnqsMod == 0) {
      I->qsrc = I->qdrn - I->qgso;
      if (I->irgateMod == 3) {
        I->gcgmgmb = (I->cgdo + I->cgso + p->cgbo) * ag0;
        I->gcgmdb = -I->cgdo * ag0;
        I->gcgmsb = -I->cgso * ag0;
        I->gcgmbb = -p->cgbo * ag0;
        I->gcdgmb = I->gcgmdb;
        I->gcsgmb = I->gcgmsb;
        I->gcbgmb = I->gcgmbb;
        I->gcggb = I->cggb * ag0;
        I->gcgdb = I->cgsb * ag0;
        I->gcgsb = I->cgdb * ag0;
        I->gcgbb = -(I->gcggb + I->gcgdb + I->gcgsb);
        I->gcdgb = -(I->cggb + I->cbgb + I->cdgb) * ag0;
        I->gcsgb = I->cdgb * ag0;
        I->gcbgb = I->cbgb * ag0;
        I->qgmb = p->cgbo * I->vgmb;
        I->qgmid = I->qgdo + I->qgso + I->qgmb;
        I->qbulk -= I->qgmb;
        I->qdrn = -(I->qgate + I->qgmid + I->qbulk + I->qsrc);
      } else {
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
      }
      I->gcddb = (I->capbd + I->cgdo - (I->cgsb + I->cbsb + I->cdsb)) * ag0;
      I->gcdsb = -(I->cgdb + I->cbdb + I->cddb) * ag0;
      I->gcsdb = I->cdsb * ag0;
      I->gcssb = (I->cddb + I->capbs + I->cgso) * ag0;
      if (!I->irbodyMod) {
        I->gcdbb = -(I->gcdgb + I->gcddb + I->gcdsb + I->gcdgmb);
        I->gcsbb = -(I->gcsgb + I->gcsdb + I->gcssb + I->gcsgmb);
        I->gcbdb = (I->cbsb - I->capbd) * ag0;
        I->gcbsb = (I->cbdb - I->capbs) * ag0;
        I->gcdbdb = 0.0;
        I->gcsbsb = 0.0;
      } else {
        I->gcdbb = -(I->gcdgb + I->gcddb + I->gcdsb + I->gcdgmb) + I->capbd * ag0;
        I->gcsbb = -(I->cddb + I->cdgb + I->cdsb) * ag0;
        I->gcbdb = I->cbsb * ag0;
        I->gcbsb = I->cbdb * ag0;
        I->gcdbdb = -I->capbd * ag0;
        I->gcsbsb = -I->capbs * ag0;
      }
      I->gcbbb = -(I->gcbgb + I->gcbdb + I->gcbsb + I->gcbgmb);
      I->ggtg = I->ggtd = I->ggtb = I->ggts = 0.0;
      I->sxpart = 0.4;
      I->dxpart = 0.6;
      I->ddxpart_dVd = I->ddxpart_dVg = I->ddxpart_dVb = I->ddxpart_dVs = 0.0;
      I->dsxpart_dVd = I->dsxpart_dVg = I->dsxpart_dVb = I->dsxpart_dVs = 0.0;
    } else {
      I->qcheq = I->qchqs;
      I->CoxWL = m->coxe * p->weffCV * I->nf * p->leffCV;
      T0 = I->qdef * I->ScalingFactor / I->CoxWL;
      I->ggtg = I->gtg = T0 * I->gcrgg;
      I->ggts = I->gts = T0 * I->gcrgd;
      I->ggtd = I->gtd = T0 * I->gcrgs;
      I->ggtb = I->gtb = T0 * I->gcrgb;
      I->gqdef = I->ScalingFactor * ag0;
      I->gcqgb = I->cqgb * ag0;
      I->gcqdb = I->cqsb * ag0;
      I->gcqsb = I->cqdb * ag0;
      I->gcqbb = I->cqbb * ag0;
      if (fabs(I->qcheq) <= 1.0e-5 * I->CoxWL) {
        if (M(xpart) < 0.5) I->sxpart = 0.4;
        else if (M(xpart) > 0.5) I->sxpart = 0.0;
        else I->sxpart = 0.5;
        I->dsxpart_dVd = I->dsxpart_dVg = I->dsxpart_dVb = I->dsxpart_dVs = 0.0;
      } else {
        I->sxpart = I->qdrn / I->qcheq;
        Css = I->cddb;
        Cds = -(I->cgdb + I->cddb + I->cbdb);
        I->dsxpart_dVs = (Css - I->sxpart * (Css + Cds)) / I->qcheq;
        Csg = I->cdgb;
        Cdg = -(I->cggb + I->cdgb + I->cbgb);
        I->dsxpart_dVg = (Csg - I->sxpart * (Csg + Cdg)) / I->qcheq;
        Csd = I->cdsb;
        Cdd = -(I->cgsb + I->cdsb + I->cbsb);
        I->dsxpart_dVd = (Csd - I->sxpart * (Csd + Cdd)) / I->qcheq;
        I->dsxpart_dVb = -(I->dsxpart_dVd + I->dsxpart_dVg + I->dsxpart_dVs);
      }
      I->dxpart = 1.0 - I->sxpart;
      I->ddxpart_dVd = -I->dsxpart_dVd;
      I->ddxpart_dVg = -I->dsxpart_dVg;
      I->ddxpart_dVs = -I->dsxpart_dVs;
      I->ddxpart_dVb = -(I->ddxpart_dVd + I->ddxpart_dVg + I->ddxpart_dVs);
      if (I->irgateMod == 3) {
        I->gcgmgmb = (I->cgdo + I->cgso + p->cgbo) * ag0;
        I->gcgmdb = -I->cgdo * ag0;
        I->gcgmsb = -I->cgso * ag0;
        I->gcgmbb = -p->cgbo * ag0;
        I->gcdgmb = I->gcgmdb;
        I->gcsgmb = I->gcgmsb;
        I->gcbgmb = I->gcgmbb;
        I->gcdgb = I->gcsgb = I->gcbgb = 0.0;
        I->gcggb = I->gcgdb = I->gcgsb = I->gcgbb = 0.0;
        I->qgmb = p->cgbo * I->vgmb;
        I->qgmid = I->qgdo + I->qgso + I->qgmb;
        I->qgate = 0.0;
        I->qbulk = -I->qgmb;
        I->qdrn = -I->qgdo;
        I->qsrc = -I->qgso;
      } else {
        I->gcggb = (I->cgdo + I->cgso + p->cgbo) * ag0;
        I->gcgdb = -I->cgdo * ag0;
        I->gcgsb = -I->cgso * ag0;
        I->gcgbb = -p->cgbo * ag0;
        I->gcdgb = I->gcgdb;
        I->gcsgb = I->gcgsb;
        I->gcbgb = I->gcgbb;
        I->gcdgmb = I->gcsgmb = I->gcbgmb = 0.0;
        I->qgb = p->cgbo * I->vgb;
        I->qgate = I->qgdo + I->qgso + I->qgb;
        I->qbulk = -I->qgb;
        I->qdrn = -I->qgdo;
        I->qsrc = -I->qgso;
      }
      I->gcddb = (I->capbd + I->cgdo) * ag0;
      I->gcdsb = I->gcsdb = 0.0;
      I->gcssb = (I->capbs + I->cgso) * ag0;
      if (!I->irbodyMod) {
        I->gcdbb = -(I->gcdgb + I->gcddb + I->gcdgmb);
        I->gcsbb = -(I->gcsgb + I->gcssb + I->gcsgmb);
        I->gcbdb = -I->capbd * ag0;
        I->gcbsb = -I->capbs * ag0;
        I->gcdbdb = 0.0;
        I->gcsbsb = 0.0;
      } else {
        I->gcdbb = I->gcsbb = I->gcbdb = I->gcbsb = 0.0;
        I->gcdbdb = -I->capbd * ag0;
        I->gcsbsb = -I->capbs * ag0;
      }
      I->gcbbb = -(I->gcbdb + I->gcbgb + I->gcbsb + I->gcbgmb);
    }
  }
}

/* setupCapacitors_newDAE (:12357-12606): the dQ/dx capacitances.  The
 * gate-mid values are set only with rgateMod 3 and the NQS ones only with
 * trnqsMod; the loads read them only then. */
void b4_setup_capacitors_new(b4inst* I) {
  const b4size* p = I->paramPtr;
  if (I->mode > 0) {
    if (I->itrnqsMod == 0) {
      if (I->irgateMod == 3) {
        I->CAPcgmgmb = (I->cgdo + I->cgso + p->cgbo);
        I->CAPcgmdb = -I->cgdo;
        I->CAPcgmsb = -I->cgso;
        I->CAPcgmbb = -p->cgbo;
        I->CAPcdgmb = I->CAPcgmdb;
        I->CAPcsgmb = I->CAPcgmsb;
        I->CAPcbgmb = I->CAPcgmbb;
        I->CAPcggb = I->cggb;
        I->CAPcgdb = I->cgdb;
        I->CAPcgsb = I->cgsb;
        I->CAPcgbb = -(I->CAPcggb + I->CAPcgdb + I->CAPcgsb);
        I->CAPcdgb = I->cdgb;
        I->CAPcsgb = -(I->cggb + I->cbgb + I->cdgb);
        I->CAPcbgb = I->cbgb;
      } else {
        I->CAPcggb = (I->cggb + I->cgdo + I->cgso + p->cgbo);
        I->CAPcgdb = (I->cgdb - I->cgdo);
        I->CAPcgsb = (I->cgsb - I->cgso);
        I->CAPcgbb = -(I->CAPcggb + I->CAPcgdb + I->CAPcgsb);
        I->CAPcdgb = (I->cdgb - I->cgdo);
        I->CAPcsgb = -(I->cggb + I->cbgb + I->cdgb + I->cgso);
        I->CAPcbgb = (I->cbgb - p->cgbo);
        I->CAPcdgmb = I->CAPcsgmb = I->CAPcbgmb = 0.0;
      }
      I->CAPcddb = (I->cddb + I->capbd + I->cgdo);
      I->CAPcdsb = I->cdsb;
      I->CAPcsdb = -(I->cgdb + I->cbdb + I->cddb);
      I->CAPcssb = (I->capbs + I->cgso - (I->cgsb + I->cbsb + I->cdsb));
      if (!I->irbodyMod) {
        I->CAPcdbb = -(I->CAPcdgb + I->CAPcddb + I->CAPcdsb + I->CAPcdgmb);
        I->CAPcsbb = -(I->CAPcsgb + I->CAPcsdb + I->CAPcssb + I->CAPcsgmb);
        I->CAPcbdb = (I->cbdb - I->capbd);
        I->CAPcbsb = (I->cbsb - I->capbs);
        I->CAPcdbdb = 0.0;
        I->CAPcsbsb = 0.0;
      } else {
        I->CAPcdbb = -(I->cddb + I->cdgb + I->cdsb);
        I->CAPcsbb = -(I->CAPcsgb + I->CAPcsdb + I->CAPcssb + I->CAPcsgmb) + I->capbs;
        I->CAPcbdb = I->cbdb;
        I->CAPcbsb = I->cbsb;
        I->CAPcdbdb = -I->capbd;
        I->CAPcsbsb = -I->capbs;
      }
      I->CAPcbbb = -(I->CAPcbdb + I->CAPcbgb + I->CAPcbsb + I->CAPcbgmb);
    } else {
      I->CAPcqgb = I->cqgb;
      I->CAPcqdb = I->cqdb;
      I->CAPcqsb = I->cqsb;
      I->CAPcqbb = I->cqbb;
      if (I->irgateMod == 3) {
        I->CAPcgmgmb = (I->cgdo + I->cgso + p->cgbo);
        I->CAPcgmdb = -I->cgdo;
        I->CAPcgmsb = -I->cgso;
        I->CAPcgmbb = -p->cgbo;
        I->CAPcdgmb = I->CAPcgmdb;
        I->CAPcsgmb = I->CAPcgmsb;
        I->CAPcbgmb = I->CAPcgmbb;
        I->CAPcdgb = I->CAPcsgb = I->CAPcbgb = 0.0;
        I->CAPcggb = I->CAPcgdb = I->CAPcgsb = I->CAPcgbb = 0.0;
      } else {
        I->CAPcggb = (I->cgdo + I->cgso + p->cgbo);
        I->CAPcgdb = -I->cgdo;
        I->CAPcgsb = -I->cgso;
        I->CAPcgbb = -p->cgbo;
        I->CAPcdgb = I->CAPcgdb;
        I->CAPcsgb = I->CAPcgsb;
        I->CAPcbgb = I->CAPcgbb;
        I->CAPcdgmb = I->CAPcsgmb = I->CAPcbgmb = 0.0;
      }
      I->CAPcddb = (I->capbd + I->cgdo);
      I->CAPcdsb = I->CAPcsdb = 0.0;
      I->CAPcssb = (I->capbs + I->cgso);
      if (!I->irbodyMod) {
        I->CAPcdbb = -(I->CAPcdgb + I->CAPcddb + I->CAPcdgmb);
        I->CAPcsbb = -(I->CAPcsgb + I->CAPcssb + I->CAPcsgmb);
        I->CAPcbdb = -I->capbd;
        I->CAPcbsb = -I->capbs;
        I->CAPcdbdb = 0.0;
        I->CAPcsbsb = 0.0;
      } else {
        I->CAPcdbb = I->CAPcsbb = I->CAPcbdb = I->CAPcbsb = 0.0;
        I->CAPcdbdb = -I->capbd;
        I->CAPcsbsb = -I->capbs;
      }
      I->CAPcbbb = -(I->CAPcbdb + I->CAPcbgb + I->CAPcbsb + I->CAPcbgmb);
    }
  } else {
    if (I->itrnqsMod == 0) {
      if (I->irgateMod == 3) {
        I->CAPcgmgmb = (I->cgdo + I->cgso + p->cgbo);
        I->CAPcgmdb = -I->cgdo;
        I->CAPcgmsb = -I->cgso;
        I->CAPcgmbb = -p->cgbo;
        I->CAPcdgmb = I->CAPcgmdb;
        I->CAPcsgmb = I->CAPcgmsb;
        I->CAPcbgmb = I->CAPcgmbb;
        I->CAPcggb = I->cggb;
        I->CAPcgdb = I->cgsb;
        I->CAPcgsb = I->cgdb;
        I->CAPcgbb = -(I->CAPcggb + I->CAPcgdb + I->CAPcgsb);
        I->CAPcdgb = -(I->cggb + I->cbgb + I->cdgb);
        I->CAPcsgb = I->cdgb;
        I->CAPcbgb = I->cbgb;
      } else {
        I->CAPcggb = (I->cggb + I->cgdo + I->cgso + p->cgbo);
        I->CAPcgdb = (I->cgsb - I->cgdo);
        I->CAPcgsb = (I->cgdb - I->cgso);
        I->CAPcgbb = -(I->CAPcggb + I->CAPcgdb + I->CAPcgsb);
        I->CAPcdgb = -(I->cggb + I->cbgb + I->cdgb + I->cgdo);
        I->CAPcsgb = (I->cdgb - I->cgso);
        I->CAPcbgb = (I->cbgb - p->cgbo);
        I->CAPcdgmb = I->CAPcsgmb = I->CAPcbgmb = 0.0;
      }
      I->CAPcddb = (I->capbd + I->cgdo - (I->cgsb + I->cbsb + I->cdsb));
      I->CAPcdsb = -(I->cgdb + I->cbdb + I->cddb);
      I->CAPcsdb = I->cdsb;
      I->CAPcssb = (I->cddb + I->capbs + I->cgso);
      if (!I->irbodyMod) {
        I->CAPcdbb = -(I->CAPcdgb + I->CAPcddb + I->CAPcdsb + I->CAPcdgmb);
        I->CAPcsbb = -(I->CAPcsgb + I->CAPcsdb + I->CAPcssb + I->CAPcsgmb);
        I->CAPcbdb = (I->cbsb - I->capbd);
        I->CAPcbsb = (I->cbdb - I->capbs);
        I->CAPcdbdb = 0.0;
        I->CAPcsbsb = 0.0;
      } else {
        I->CAPcdbb = -(I->CAPcdgb + I->CAPcddb + I->CAPcdsb + I->CAPcdgmb) + I->capbd;
        I->CAPcsbb = -(I->cddb + I->cdgb + I->cdsb);
        I->CAPcbdb = I->cbsb;
        I->CAPcbsb = I->cbdb;
        I->CAPcdbdb = -I->capbd;
        I->CAPcsbsb = -I->capbs;
      }
      I->CAPcbbb = -(I->CAPcbgb + I->CAPcbdb + I->CAPcbsb + I->CAPcbgmb);
    } else {
      I->CAPcqgb = I->cqgb;
      I->CAPcqdb = I->cqsb;
      I->CAPcqsb = I->cqdb;
      I->CAPcqbb = I->cqbb;
      if (I->irgateMod == 3) {
        I->CAPcgmgmb = (I->cgdo + I->cgso + p->cgbo);
        I->CAPcgmdb = -I->cgdo;
        I->CAPcgmsb = -I->cgso;
        I->CAPcgmbb = -p->cgbo;
        I->CAPcdgmb = I->CAPcgmdb;
        I->CAPcsgmb = I->CAPcgmsb;
        I->CAPcbgmb = I->CAPcgmbb;
        I->CAPcdgb = I->CAPcsgb = I->CAPcbgb = 0.0;
        I->CAPcggb = I->CAPcgdb = I->CAPcgsb = I->CAPcgbb = 0.0;
      } else {
        I->CAPcggb = (I->cgdo + I->cgso + p->cgbo);
        I->CAPcgdb = -I->cgdo;
        I->CAPcgsb = -I->cgso;
        I->CAPcgbb = -p->cgbo;
        I->CAPcdgb = I->CAPcgdb;
        I->CAPcsgb = I->CAPcgsb;
        I->CAPcbgb = I->CAPcgbb;
        I->CAPcdgmb = I->CAPcsgmb = I->CAPcbgmb = 0.0;
      }
      I->CAPcddb = (I->capbd + I->cgdo);
      I->CAPcdsb = I->CAPcsdb = 0.0;
      I->CAPcssb = (I->capbs + I->cgso);
      if (!I->irbodyMod) {
        I->CAPcdbb = -(I->CAPcdgb + I->CAPcddb + I->CAPcdgmb);
        I->CAPcsbb = -(I->CAPcsgb + I->CAPcssb + I->CAPcsgmb);
        I->CAPcbdb = -I->capbd;
        I->CAPcbsb = -I->capbs;
        I->CAPcdbdb = 0.0;
        I->CAPcsbsb = 0.0;
      } else {
        I->CAPcdbb = I->CAPcsbb = I->CAPcbdb = I->CAPcbsb = 0.0;
        I->CAPcdbdb = -I->capbd;
        I->CAPcsbsb = -I->capbs;
      }
      I->CAPcbbb = -(I->CAPcbdb + I->CAPcbgb + I->CAPcbsb + I->CAPcbgmb);
    }
  }
}

/* auxChargeCalculations (:12216-12347) */
static void aux_charge(b4inst* I) {
  const b4model* m = I->model;
  const b4size* p = I->paramPtr;
  double T0, T1;
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
    if (I->itrnqsMod) {
      I->CoxWL = m->coxe * p->weffCV * I->nf * p->leffCV;
      T1 = I->gcrg / I->CoxWL;
      I->gtau = T1 * I->ScalingFactor;
    } else {
      I->gtau = 0.0;
    }
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
    if (I->irgateMod == 3) {
      I->Qeqqgmid_Jdxp = 0.0;
      if (!I->origFlag)
        I->Qeqqgmid_Jdxp = +I->CAPcgmdb * (I->vbd - I->vbd_orig) + I->CAPcgmsb * (I->vbs - I->vbs_orig)
                           - I->CAPcgmgmb * (I->vgmb - I->vgmb_orig);
    } else {
      I->Qeqqgmid_Jdxp = 0.0;
    }
    if (I->irbodyMod) {
      I->Qeqqjs_Jdxp = 0.0;
      I->Qeqqjd_Jdxp = 0.0;
      if (!I->origFlag) {
        I->Qeqqjs_Jdxp = I->CAPcsbsb * (I->vbs_jct - I->vbs_jct_orig);
        I->Qeqqjd_Jdxp = I->CAPcdbdb * (I->vbd_jct - I->vbd_jct_orig);
      }
    }
    if (I->itrnqsMod) {
      T0 = I->ggtg * (I->vgb - I->vgb_orig) - I->ggtd * (I->vbd - I->vbd_orig)
           - I->ggts * (I->vbs - I->vbs_orig);
      I->Qeqqg_Jdxp += T0;
      T1 = I->qdef * I->gtau;
      I->Qeqqd_Jdxp -= I->dxpart * T0 + T1 * (I->ddxpart_dVg * (I->vgb - I->vgb_orig)
                                              - I->ddxpart_dVd * (I->vbd - I->vbd_orig)
                                              - I->ddxpart_dVs * (I->vbs - I->vbs_orig));
      I->Qqcheq_Jdxp = -(I->CAPcqgb * (I->vgb - I->vgb_orig) - I->CAPcqdb * (I->vbd - I->vbd_orig)
                         - I->CAPcqsb * (I->vbs - I->vbs_orig)) + T0;
    }
  }
}

/* setupFVectorVars (:13525-13838) */
static void setup_fvector_vars(b4inst* I) {
  const b4model* m = I->model;
  const double dt = m->dtype;
  double T0 = 0.0, tmp_gcrgd;
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
    if (M(igcMod)) {
      I->gIstotg = I->gIgsg + I->gIgcsg;
      I->gIstotd = I->gIgcsd;
      I->gIstots = I->gIgss + I->gIgcss;
      I->gIstotb = I->gIgcsb;
      I->Istoteq = dt * (I->Igs + I->Igcs);
      I->Istoteq_Jdxp = dt * (-I->gIstotg * (I->vgs - I->vgs_orig) - I->gIgcsd * (I->vds - I->vds_orig)
                              - I->gIgcsb * (I->vbs - I->vbs_orig));
      I->gIdtotg = I->gIgdg + I->gIgcdg;
      I->gIdtotd = I->gIgdd + I->gIgcdd;
      I->gIdtots = I->gIgcds;
      I->gIdtotb = I->gIgcdb;
      I->Idtoteq = dt * (I->Igd + I->Igcd);
      I->Idtoteq_Jdxp = dt * (-I->gIgdg * (I->vgd - I->vgd_orig) - I->gIgcdg * (I->vgs - I->vgs_orig)
                              - I->gIgcdd * (I->vds - I->vds_orig) - I->gIgcdb * (I->vbs - I->vbs_orig));
    } else {
      I->gIstotg = I->gIstotd = I->gIstots = I->gIstotb = I->Istoteq = 0.0;
      I->gIdtotg = I->gIdtotd = I->gIdtots = I->gIdtotb = I->Idtoteq = 0.0;
      I->Istoteq_Jdxp = 0.0;
      I->Idtoteq_Jdxp = 0.0;
    }
    if (M(igbMod)) {
      I->gIbtotg = I->gIgbg;
      I->gIbtotd = I->gIgbd;
      I->gIbtots = I->gIgbs;
      I->gIbtotb = I->gIgbb;
      I->Ibtoteq = dt * I->Igb;
      I->Ibtoteq_Jdxp = dt * (-I->gIgbg * (I->vgs - I->vgs_orig) - I->gIgbd * (I->vds - I->vds_orig)
                              - I->gIgbb * (I->vbs - I->vbs_orig));
    } else {
      I->gIbtotg = I->gIbtotd = I->gIbtots = I->gIbtotb = I->Ibtoteq = 0.0;
      I->Ibtoteq_Jdxp = 0.0;
    }
    if ((M(igcMod) != 0) || (M(igbMod) != 0)) {
      I->gIgtotg = I->gIstotg + I->gIdtotg + I->gIbtotg;
      I->gIgtotd = I->gIstotd + I->gIdtotd + I->gIbtotd;
      I->gIgtots = I->gIstots + I->gIdtots + I->gIbtots;
      I->gIgtotb = I->gIstotb + I->gIdtotb + I->gIbtotb;
      I->Igtoteq = I->Istoteq + I->Idtoteq + I->Ibtoteq;
      I->Igtoteq_Jdxp = I->Istoteq_Jdxp + I->Idtoteq_Jdxp + I->Ibtoteq_Jdxp;
    } else {
      I->gIgtotg = I->gIgtotd = I->gIgtots = I->gIgtotb = I->Igtoteq = 0.0;
      I->Igtoteq_Jdxp = 0.0;
    }
    if (I->irgateMod == 2) T0 = I->vges - I->vgs;
    else if (I->irgateMod == 3) T0 = I->vgms - I->vgs;
    if (I->irgateMod > 1) {
      I->gcrgd = I->gcrgd * T0;
      I->gcrgg = I->gcrgg * T0;
      I->gcrgs = I->gcrgs * T0;
      I->gcrgb = I->gcrgb * T0;
      I->ceqgcrg = 0.0;
      I->ceqgcrg_Jdxp = -(I->gcrgd * (I->vds - I->vds_orig) + I->gcrgg * (I->vgs - I->vgs_orig)
                          + I->gcrgb * (I->vbs - I->vbs_orig));
      I->gcrgg -= I->gcrg;
    } else {
      I->ceqgcrg = I->gcrg = I->gcrgd = I->gcrgg = I->gcrgs = I->gcrgb = 0.0;
      I->ceqgcrg_Jdxp = 0.0;
    }
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
    if (M(igcMod)) {
      I->gIstotg = I->gIgsg + I->gIgcdg;
      I->gIstotd = I->gIgcds;
      I->gIstots = I->gIgss + I->gIgcdd;
      I->gIstotb = I->gIgcdb;
      I->Istoteq = dt * (I->Igs + I->Igcd);
      I->Istoteq_Jdxp = dt * (-I->gIgsg * (I->vgs - I->vgs_orig) - I->gIgcdg * (I->vgd - I->vgd_orig)
                              + I->gIgcdd * (I->vds - I->vds_orig) - I->gIgcdb * (I->vbd - I->vbd_orig));
      I->gIdtotg = I->gIgdg + I->gIgcsg;
      I->gIdtotd = I->gIgdd + I->gIgcss;
      I->gIdtots = I->gIgcsd;
      I->gIdtotb = I->gIgcsb;
      I->Idtoteq = dt * (I->Igd + I->Igcs);
      I->Idtoteq_Jdxp = dt * (-(I->gIgdg + I->gIgcsg) * (I->vgd - I->vgd_orig)
                              + I->gIgcsd * (I->vds - I->vds_orig) - I->gIgcsb * (I->vbd - I->vbd_orig));
    } else {
      I->gIstotg = I->gIstotd = I->gIstots = I->gIstotb = I->Istoteq = 0.0;
      I->gIdtotg = I->gIdtotd = I->gIdtots = I->gIdtotb = I->Idtoteq = 0.0;
      I->Istoteq_Jdxp = 0.0;
      I->Idtoteq_Jdxp = 0.0;
    }
    if (M(igbMod)) {
      I->gIbtotg = I->gIgbg;
      I->gIbtotd = I->gIgbs;
      I->gIbtots = I->gIgbd;
      I->gIbtotb = I->gIgbb;
      I->Ibtoteq = dt * I->Igb;
      I->Ibtoteq_Jdxp = dt * (-I->gIgbg * (I->vgd - I->vgd_orig) + I->gIgbd * (I->vds - I->vds_orig)
                              - I->gIgbb * (I->vbd - I->vbd_orig));
    } else {
      I->gIbtotg = I->gIbtotd = I->gIbtots = I->gIbtotb = I->Ibtoteq = 0.0;
      I->Ibtoteq_Jdxp = 0.0;
    }
    if ((M(igcMod) != 0) || (M(igbMod) != 0)) {
      I->gIgtotg = I->gIstotg + I->gIdtotg + I->gIbtotg;
      I->gIgtotd = I->gIstotd + I->gIdtotd + I->gIbtotd;
      I->gIgtots = I->gIstots + I->gIdtots + I->gIbtots;
      I->gIgtotb = I->gIstotb + I->gIdtotb + I->gIbtotb;
      I->Igtoteq = I->Istoteq + I->Idtoteq + I->Ibtoteq;
      I->Igtoteq_Jdxp = I->Istoteq_Jdxp + I->Idtoteq_Jdxp + I->Ibtoteq_Jdxp;
    } else {
      I->gIgtotg = I->gIgtotd = I->gIgtots = I->gIgtotb = I->Igtoteq = 0.0;
      I->Igtoteq_Jdxp = 0.0;
    }
    if (I->irgateMod == 2) T0 = I->vges - I->vgs;
    else if (I->irgateMod == 3) T0 = I->vgms - I->vgs;
    if (I->irgateMod > 1) {
      tmp_gcrgd = I->gcrgd;
      I->gcrgd = I->gcrgs * T0;
      I->gcrgg = I->gcrgg * T0;
      I->gcrgs = tmp_gcrgd * T0;
      I->gcrgb = I->gcrgb * T0;
      I->ceqgcrg = 0.0;
      I->ceqgcrg_Jdxp = -(I->gcrgg * (I->vgd - I->vgd_orig) - I->gcrgs * (I->vds - I->vds_orig)
                          + I->gcrgb * (I->vbd - I->vbd_orig));
      I->gcrgg -= I->gcrg;
    } else {
      /* the reverse branch leaves ceqgcrg_Jdxp at its zero */
      I->ceqgcrg = I->gcrg = I->gcrgd = I->gcrgg = I->gcrgs = I->gcrgb = 0.0;
    }
  }
  if (M(rdsMod) == 1) {
    I->ceqgstot = 0.0;
    I->ceqgstot_Jdxp = dt * (I->gstotd * (I->vds - I->vds_orig) + I->gstotg * (I->vgs - I->vgs_orig)
                             + I->gstotb * (I->vbs - I->vbs_orig));
    I->gstots = I->gstots - I->gstot;
    I->ceqgdtot = 0.0;
    I->ceqgdtot_Jdxp = -dt * (I->gdtotd * (I->vds - I->vds_orig) + I->gdtotg * (I->vgs - I->vgs_orig)
                              + I->gdtotb * (I->vbs - I->vbs_orig));
    I->gdtotd = I->gdtotd - I->gdtot;
  } else {
    I->gstot = I->gstotd = I->gstotg = I->gstots = I->gstotb = I->ceqgstot = 0.0;
    I->gdtot = I->gdtotd = I->gdtotg = I->gdtots = I->gdtotb = I->ceqgdtot = 0.0;
    I->ceqgstot_Jdxp = 0.0;
    I->ceqgdtot_Jdxp = 0.0;
  }
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

/* registerLIDs (:4996-5128), registerStateLIDs (:5200-5257),
 * registerStoreLIDs (:5267-5290) */
static void register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes, const int32_t* sta,
                          const int32_t* sto) {
  b4inst* I = &((b4inst*)g->inst)[i];
  int loc = 4, ls = 3;
  I->li_Drain = nodes[0];
  I->li_GateExt = nodes[1];
  I->li_Source = nodes[2];
  I->li_Body = nodes[3];
  I->li_DrainPrime = I->drainMOSFET_B4Exists ? nodes[loc++] : I->li_Drain;
  I->li_SourcePrime = I->sourceMOSFET_B4Exists ? nodes[loc++] : I->li_Source;
  I->li_GatePrime = (I->irgateMod > 0) ? nodes[loc++] : I->li_GateExt;
  I->li_GateMid = (I->irgateMod == 3) ? nodes[loc++] : I->li_GateExt;
  if (I->irbodyMod) {
    I->li_BodyPrime = nodes[loc++];
    I->li_SourceBody = nodes[loc++];
    I->li_DrainBody = nodes[loc++];
  } else {
    I->li_BodyPrime = I->li_Body;
    I->li_SourceBody = I->li_Body;
    I->li_DrainBody = I->li_Body;
  }
  I->li_Charge = I->itrnqsMod ? nodes[loc++] : -1;
  I->li_Ibs = I->icVBS_given ? nodes[loc++] : -1;
  I->li_Ids = I->icVDS_given ? nodes[loc++] : -1;
  I->li_Igs = I->icVGS_given ? nodes[loc++] : -1;
  I->li_state_qb = sta ? sta[0] : -1;
  I->li_state_qg = sta ? sta[1] : -1;
  I->li_state_qd = sta ? sta[2] : -1;
  I->li_state_qgmid = (I->irgateMod == 3 && sta) ? sta[ls++] : -1;
  if (I->irbodyMod && sta) {
    I->li_state_qbs = sta[ls++];
    I->li_state_qbd = sta[ls++];
  } else {
    I->li_state_qbs = I->li_state_qbd = -1;
  }
  if (I->itrnqsMod && sta) {
    I->li_state_qcheq = sta[ls++];
    I->li_state_qcdump = sta[ls++];
  } else {
    I->li_state_qcheq = I->li_state_qcdump = -1;
  }
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

/* Which pairs this instance's loads address (its options' branches of
 * loadDAEMatrices; the IC pairs in both the DC and the transient form). */
static void pairs_used(const b4inst* I, char use[B4_NPAIR]) {
  static const int base[] = {B4_GPgp, B4_GPdp, B4_GPsp, B4_GPbp, B4_DPdp, B4_DPd, B4_DPgp, B4_DPsp,
                             B4_DPbp, B4_Ddp, B4_Dd, B4_SPdp, B4_SPgp, B4_SPsp, B4_SPs, B4_SPbp,
                             B4_Ssp, B4_Ss, B4_BPdp, B4_BPgp, B4_BPsp, B4_BPbp};
  static const int rg1[] = {B4_GEge, B4_GEgp, B4_GPge};
  static const int rg2[] = {B4_GEge, B4_GEgp, B4_GEdp, B4_GEsp, B4_GEbp, B4_GPge};
  static const int rg3[] = {B4_GEge, B4_GEgm, B4_GMge, B4_GMgm, B4_GMdp, B4_GMgp, B4_GMsp, B4_GMbp,
                            B4_GPgm, B4_DPgm, B4_SPgm, B4_BPgm};
  static const int rds[] = {B4_Dgp, B4_Dsp, B4_Dbp, B4_Sdp, B4_Sgp, B4_Sbp};
  static const int rbd[] = {B4_DPdb, B4_SPsb, B4_DBdp, B4_DBdb, B4_DBbp, B4_DBb, B4_BPdb, B4_BPb,
                            B4_BPsb, B4_SBsp, B4_SBbp, B4_SBb, B4_SBsb, B4_Bdb, B4_Bbp, B4_Bsb, B4_Bb};
  static const int nqs[] = {B4_Qq, B4_Qgp, B4_Qdp, B4_Qsp, B4_Qbp, B4_DPq, B4_SPq, B4_GPq};
  static const int ibs[] = {B4_Bibs, B4_Sibs, B4_IBSb, B4_IBSs, B4_IBSibs};
  static const int ids[] = {B4_Dids, B4_Sids, B4_IDSd, B4_IDSs, B4_IDSids};
  static const int igs[] = {B4_GEigs, B4_Sigs, B4_IGSg, B4_IGSs, B4_IGSigs};
  const b4model* m = I->model;
  size_t k;
#define USE(a) for (k = 0; k < sizeof(a) / sizeof(a[0]); ++k) use[a[k]] = 1
  memset(use, 0, B4_NPAIR);
  USE(base);
  if (I->irgateMod == 1) USE(rg1);
  else if (I->irgateMod == 2) USE(rg2);
  else if (I->irgateMod == 3) USE(rg3);
  if (M(rdsMod)) USE(rds);
  if (I->irbodyMod) USE(rbd);
  if (I->itrnqsMod) USE(nqs);
  if (I->icVBS_given) USE(ibs);
  if (I->icVDS_given) USE(ids);
  if (I->icVGS_given) USE(igs);
#undef USE
}

static void pair_nodes(const b4inst* I, int k, int32_t* r, int32_t* c) {
  switch (k) {
#define B4_PAIR_CASE(name, rn, cn) case B4_##name: *r = I->li_##rn; *c = I->li_##cn; break;
    B4_PAIRS(B4_PAIR_CASE)
#undef B4_PAIR_CASE
    default: *r = *c = -1;
  }
}

static int64_t jac_pairs(struct or_group_s* g, int64_t i, int32_t* r, int32_t* c) {
  const b4inst* I = &((b4inst*)g->inst)[i];
  char use[B4_NPAIR];
  int k;
  int64_t n = 0;
  pairs_used(I, use);
  for (k = 0; k < B4_NPAIR; ++k) {
    int32_t pr, pc;
    if (!use[k]) continue;
    pair_nodes(I, k, &pr, &pc);
    if (pr < 0 || pc < 0) continue;
    if (r) { r[n] = pr; c[n] = pc; }
    ++n;
  }
  return n;
}

static int register_jac(struct or_group_s* g, int64_t i) {
  b4inst* I = &((b4inst*)g->inst)[i];
  char use[B4_NPAIR];
  int k;
  pairs_used(I, use);
  for (k = 0; k < B4_NPAIR; ++k) {
    int32_t pr, pc;
    I->sp[k] = -1;
    if (!use[k]) continue;
    pair_nodes(I, k, &pr, &pc);
    I->sp[k] = or_slot(g, pr, pc);
    if (I->sp[k] < -1) return XA_ERR_ARG;
  }
  return XA_OK;
}

/* Master::updateState (:16180-16293) */
static void put_state(double* s, int lid, double val) { if (s && lid >= 0) s[lid] = val; }

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
    if (!I->irbodyMod) I->qb = I->qbulk + I->qbd + I->qbs;
    else I->qb = I->qbulk;
    put_state(sta, I->li_state_qg, I->qg);
    put_state(sta, I->li_state_qd, I->qd);
    put_state(sta, I->li_state_qb, I->qb);
    if (I->irgateMod == 3) put_state(sta, I->li_state_qgmid, I->qgmid);
    if (I->irbodyMod) {
      put_state(sta, I->li_state_qbs, I->qbs);
      put_state(sta, I->li_state_qbd, I->qbd);
    }
    if (I->itrnqsMod) {
      put_state(sta, I->li_state_qcheq, I->qcheq);
      put_state(sta, I->li_state_qcdump, I->qdef * I->ScalingFactor);
    }
    if (!(ss->dcopFlag) && ss->initTranFlag && ss->newtonIter == 0 && v->currSta) {
      double* cs = v->currSta;
      put_state(cs, I->li_state_qg, I->qgate);
      put_state(cs, I->li_state_qd, I->qdrn - I->qbd);
      if (!I->irbodyMod) put_state(cs, I->li_state_qb, I->qbulk + I->qbd + I->qbs);
      else put_state(cs, I->li_state_qb, I->qbulk);
      if (I->irgateMod == 3) put_state(cs, I->li_state_qgmid, I->qgmid);
      if (I->irbodyMod) {
        put_state(cs, I->li_state_qbs, I->qbs);
        put_state(cs, I->li_state_qbd, I->qbd);
      }
      if (I->itrnqsMod) {
        put_state(cs, I->li_state_qcheq, I->qcheq);
        put_state(cs, I->li_state_qcdump, I->qdef * I->ScalingFactor);
      }
    }
  }
  return ok ? XA_OK : XA_ERR_DEVICE;
}

static double solv(const double* x, int lid) { return lid >= 0 ? x[lid] : 0.0; }

/* Master::loadDAEVectors (:16303-16567) */
static int load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  b4inst* in = (b4inst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    b4inst* mi = &in[i];
    const b4model* m = mi->model;
    const double Mp = mi->numberParallel;
    double Qeqqg = 0.0, Qeqqb = 0.0, Qeqqd = 0.0, Qeqqgmid = 0.0, Qeqqjs = 0.0, Qeqqjd = 0.0;
    double Qqdef = 0.0, Qqcheq = 0.0, coef;
    setup_fvector_vars(mi);
    if (v->f) {
      double* f = v->f;
      OR_ADD(f, mi->li_DrainPrime, -(mi->ceqjd - mi->ceqbd - mi->ceqdrn + mi->Idtoteq) * Mp);
      OR_SUB(f, mi->li_GatePrime, -(-mi->ceqgcrg + mi->Igtoteq) * Mp);
      if (mi->irgateMod == 1) {
        OR_SUB(f, mi->li_GateExt, -(mi->Igate) * Mp);
        OR_ADD(f, mi->li_GatePrime, -(mi->Igate) * Mp);
      } else if (mi->irgateMod == 2) {
        OR_SUB(f, mi->li_GateExt, -(mi->Igate + mi->ceqgcrg) * Mp);
        OR_ADD(f, mi->li_GatePrime, -(mi->Igate) * Mp);
      } else if (mi->irgateMod == 3) {
        OR_SUB(f, mi->li_GateExt, -(mi->Igate) * Mp);
        OR_SUB(f, mi->li_GateMid, -(mi->IgateMid - mi->Igate + mi->ceqgcrg) * Mp);
        OR_ADD(f, mi->li_GatePrime, -(mi->IgateMid) * Mp);
      }
      if (!mi->irbodyMod) {
        OR_ADD(f, mi->li_BodyPrime, -(mi->ceqbd + mi->ceqbs - mi->ceqjd - mi->ceqjs + mi->Ibtoteq) * Mp);
        OR_ADD(f, mi->li_SourcePrime, -(mi->ceqdrn - mi->ceqbs + mi->ceqjs + mi->Istoteq) * Mp);
      } else {
        OR_SUB(f, mi->li_DrainBody, -(mi->ceqjd + mi->Idbb + mi->Idbbp) * Mp);
        OR_ADD(f, mi->li_BodyPrime, -(mi->ceqbd + mi->ceqbs + mi->Ibtoteq + mi->Idbbp + mi->Isbbp
                                      - mi->Ibpb) * Mp);
        OR_ADD(f, mi->li_Body, -(mi->Isbb + mi->Idbb + mi->Ibpb) * Mp);
        OR_SUB(f, mi->li_SourceBody, -(mi->ceqjs + mi->Isbb + mi->Isbbp) * Mp);
        OR_ADD(f, mi->li_SourcePrime, -(mi->ceqdrn - mi->ceqbs + mi->ceqjs + mi->Istoteq) * Mp);
      }
      if (M(rdsMod)) {
        OR_ADD(f, mi->li_Drain, -(-mi->ceqgdtot) * Mp);
        OR_ADD(f, mi->li_Source, -(mi->ceqgstot) * Mp);
        OR_ADD(f, mi->li_DrainPrime, -(mi->ceqgdtot) * Mp);
        OR_ADD(f, mi->li_SourcePrime, -(-mi->ceqgstot) * Mp);
      }
      if (mi->drainMOSFET_B4Exists) {
        OR_ADD(f, mi->li_Drain, -(-mi->Idrain) * Mp);
        OR_ADD(f, mi->li_DrainPrime, -(mi->Idrain) * Mp);
      }
      if (mi->sourceMOSFET_B4Exists) {
        OR_ADD(f, mi->li_Source, -(-mi->Isource) * Mp);
        OR_ADD(f, mi->li_SourcePrime, -(+mi->Isource) * Mp);
      }
      /* initial conditions (:16382-16411) */
      if (ss->dcopFlag && mi->icVBS_given) {
        coef = solv(v->nextSol, mi->li_Ibs);
        OR_ADD(f, mi->li_Body, coef);
        OR_ADD(f, mi->li_Source, -coef);
        OR_ADD(f, mi->li_Ibs, (solv(v->nextSol, mi->li_Body) - solv(v->nextSol, mi->li_Source) - mi->icVBS));
      }
      if (ss->dcopFlag && mi->icVDS_given) {
        coef = solv(v->nextSol, mi->li_Ids);
        OR_ADD(f, mi->li_Drain, coef);
        OR_ADD(f, mi->li_Source, -coef);
        OR_ADD(f, mi->li_Ids, (solv(v->nextSol, mi->li_Drain) - solv(v->nextSol, mi->li_Source) - mi->icVDS));
      }
      if (ss->dcopFlag && mi->icVGS_given) {
        coef = solv(v->nextSol, mi->li_Igs);
        OR_ADD(f, mi->li_GateExt, coef);
        OR_ADD(f, mi->li_Source, -coef);
        OR_ADD(f, mi->li_Igs, (solv(v->nextSol, mi->li_GateExt) - solv(v->nextSol, mi->li_Source) - mi->icVGS));
      }
    }
    if (g->opt.voltageLimiterFlag && !mi->origFlag && v->dFdxdVp) {
      double* f = v->dFdxdVp;
      OR_ADD(f, mi->li_DrainPrime, (mi->ceqjd_Jdxp - mi->ceqbd_Jdxp - mi->ceqdrn_Jdxp + mi->Idtoteq_Jdxp) * Mp);
      OR_SUB(f, mi->li_GatePrime, (-mi->ceqgcrg_Jdxp + mi->Igtoteq_Jdxp) * Mp);
      if (mi->irgateMod == 1) {
        OR_SUB(f, mi->li_GateExt, (mi->Igate_Jdxp) * Mp);
        OR_ADD(f, mi->li_GatePrime, (mi->Igate_Jdxp) * Mp);
      } else if (mi->irgateMod == 2) {
        OR_SUB(f, mi->li_GateExt, (mi->Igate_Jdxp + mi->ceqgcrg_Jdxp) * Mp);
        OR_ADD(f, mi->li_GatePrime, (mi->Igate_Jdxp) * Mp);
      } else if (mi->irgateMod == 3) {
        OR_SUB(f, mi->li_GateExt, (mi->Igate_Jdxp) * Mp);
        OR_SUB(f, mi->li_GateMid, (mi->IgateMid_Jdxp - mi->Igate_Jdxp + mi->ceqgcrg_Jdxp) * Mp);
        OR_ADD(f, mi->li_GatePrime, (mi->IgateMid_Jdxp) * Mp);
      }
      if (!mi->irbodyMod) {
        OR_ADD(f, mi->li_BodyPrime,
               (mi->ceqbd_Jdxp + mi->ceqbs_Jdxp - mi->ceqjd_Jdxp - mi->ceqjs_Jdxp + mi->Ibtoteq_Jdxp) * Mp);
        OR_ADD(f, mi->li_SourcePrime,
               (mi->ceqdrn_Jdxp - mi->ceqbs_Jdxp + mi->ceqjs_Jdxp + mi->Istoteq_Jdxp) * Mp);
      } else {
        OR_SUB(f, mi->li_DrainBody, (mi->ceqjd_Jdxp) * Mp);
        OR_ADD(f, mi->li_BodyPrime, (mi->ceqbd_Jdxp + mi->ceqbs_Jdxp + mi->Ibtoteq_Jdxp) * Mp);
        OR_SUB(f, mi->li_SourceBody, (mi->ceqjs_Jdxp) * Mp);
        OR_ADD(f, mi->li_SourcePrime,
               (mi->ceqdrn_Jdxp - mi->ceqbs_Jdxp + mi->ceqjs_Jdxp + mi->Istoteq_Jdxp) * Mp);
      }
      if (M(rdsMod)) {
        OR_SUB(f, mi->li_Drain, (mi->ceqgdtot_Jdxp) * Mp);
        OR_ADD(f, mi->li_Source, (mi->ceqgstot_Jdxp) * Mp);
        OR_ADD(f, mi->li_DrainPrime, (mi->ceqgdtot_Jdxp) * Mp);
        OR_SUB(f, mi->li_SourcePrime, (mi->ceqgstot_Jdxp) * Mp);
      }
    }
    aux_charge(mi);
    if (m->dtype > 0) {
      Qeqqg = mi->qg;
      Qeqqd = mi->qd;
      Qeqqb = mi->qb;
      if (mi->itrnqsMod) {
        Qqdef = mi->qdef;
        Qqcheq = mi->qcheq;
      }
      if (mi->irbodyMod) {
        Qeqqjs = mi->qbs;
        Qeqqjd = mi->qbd;
      }
      if (mi->irgateMod == 3) Qeqqgmid = mi->qgmid;
    } else {
      Qeqqg = -mi->qg;
      Qeqqd = -mi->qd;
      Qeqqb = -mi->qb;
      if (mi->itrnqsMod) {
        Qqdef = -mi->qdef;
        Qqcheq = -mi->qcheq;
      }
      if (mi->irbodyMod) {
        Qeqqjs = -mi->qbs;
        Qeqqjd = -mi->qbd;
      }
      if (mi->irgateMod == 3) Qeqqgmid = -mi->qgmid;
    }
    if (v->q) {
      double* q = v->q;
      OR_ADD(q, mi->li_DrainPrime, -(-Qeqqd) * Mp);
      OR_SUB(q, mi->li_GatePrime, -(Qeqqg) * Mp);
      if (mi->irgateMod == 3) OR_SUB(q, mi->li_GateMid, -(+Qeqqgmid) * Mp);
      if (!mi->irbodyMod) {
        OR_ADD(q, mi->li_BodyPrime, -(-Qeqqb) * Mp);
        OR_ADD(q, mi->li_SourcePrime, -(+Qeqqg + Qeqqb + Qeqqd + Qeqqgmid) * Mp);
      } else {
        OR_SUB(q, mi->li_DrainBody, -(Qeqqjd) * Mp);
        OR_ADD(q, mi->li_BodyPrime, -(-Qeqqb) * Mp);
        OR_SUB(q, mi->li_SourceBody, -(Qeqqjs) * Mp);
        OR_ADD(q, mi->li_SourcePrime, -(Qeqqd + Qeqqg + Qeqqb + Qeqqjd + Qeqqjs + Qeqqgmid) * Mp);
      }
      if (mi->itrnqsMod) OR_ADD(q, mi->li_Charge, -(Qqcheq - Qqdef) * Mp);
    }
    if (g->opt.voltageLimiterFlag && !mi->origFlag && v->dQdxdVp) {
      double* q = v->dQdxdVp;
      OR_ADD(q, mi->li_DrainPrime, (-mi->Qeqqd_Jdxp) * Mp);
      OR_SUB(q, mi->li_GatePrime, (mi->Qeqqg_Jdxp) * Mp);
      if (mi->irgateMod == 3) OR_SUB(q, mi->li_GateMid, (+mi->Qeqqgmid_Jdxp) * Mp);
      if (!mi->irbodyMod) {
        OR_ADD(q, mi->li_BodyPrime, (-mi->Qeqqb_Jdxp) * Mp);
        OR_ADD(q, mi->li_SourcePrime,
               (+mi->Qeqqg_Jdxp + mi->Qeqqb_Jdxp + mi->Qeqqd_Jdxp + mi->Qeqqgmid_Jdxp) * Mp);
      } else {
        OR_SUB(q, mi->li_DrainBody, (mi->Qeqqjd_Jdxp) * Mp);
        OR_ADD(q, mi->li_BodyPrime, (-mi->Qeqqb_Jdxp) * Mp);
        OR_SUB(q, mi->li_SourceBody, (+mi->Qeqqjs_Jdxp) * Mp);
        OR_ADD(q, mi->li_SourcePrime, (+mi->Qeqqd_Jdxp + mi->Qeqqg_Jdxp + mi->Qeqqb_Jdxp + mi->Qeqqjd_Jdxp
                                       + mi->Qeqqjs_Jdxp + mi->Qeqqgmid_Jdxp) * Mp);
      }
      if (mi->itrnqsMod) OR_ADD(q, mi->li_Charge, (mi->Qqcheq_Jdxp) * Mp);
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
    const b4model* m = mi->model;
    const double Mp = mi->numberParallel;
    const int64_t* sp = mi->sp;
    double T1;
    double* F = v->dFdx;
    double* Q = v->dQdx;
    if (!mi->irbodyMod) {
      mi->gjbd = mi->gbd;
      mi->gjbs = mi->gbs;
    } else {
      mi->gjbd = mi->gjbs = 0.0;
    }
    if (!M(rdsMod)) {
      mi->gdpr = mi->drainConductance;
      mi->gspr = mi->sourceConductance;
    } else {
      mi->gdpr = mi->gspr = 0.0;
    }
    mi->geltd = mi->grgeltd;
    T1 = mi->qdef * mi->gtau;
#define S_(x) sp[B4_##x]
    if (F) {
      if (mi->irgateMod == 1) {
        OR_ADD(F, S_(GEge), (mi->geltd) * Mp);
        OR_SUB(F, S_(GEgp), (mi->geltd) * Mp);
        OR_SUB(F, S_(GPge), (mi->geltd) * Mp);
        OR_ADD(F, S_(GPgp), (+mi->geltd - mi->ggtg + mi->gIgtotg) * Mp);
        OR_ADD(F, S_(GPdp), (-mi->ggtd + mi->gIgtotd) * Mp);
        OR_ADD(F, S_(GPsp), (-mi->ggts + mi->gIgtots) * Mp);
        OR_ADD(F, S_(GPbp), (-mi->ggtb + mi->gIgtotb) * Mp);
      } else if (mi->irgateMod == 2) {
        OR_ADD(F, S_(GEge), (mi->gcrg) * Mp);
        OR_ADD(F, S_(GEgp), (mi->gcrgg) * Mp);
        OR_ADD(F, S_(GEdp), (mi->gcrgd) * Mp);
        OR_ADD(F, S_(GEsp), (mi->gcrgs) * Mp);
        OR_ADD(F, S_(GEbp), (mi->gcrgb) * Mp);
        OR_SUB(F, S_(GPge), (mi->gcrg) * Mp);
        OR_ADD(F, S_(GPgp), (-mi->gcrgg - mi->ggtg + mi->gIgtotg) * Mp);
        OR_ADD(F, S_(GPdp), (-mi->gcrgd - mi->ggtd + mi->gIgtotd) * Mp);
        OR_ADD(F, S_(GPsp), (-mi->gcrgs - mi->ggts + mi->gIgtots) * Mp);
        OR_ADD(F, S_(GPbp), (-mi->gcrgb - mi->ggtb + mi->gIgtotb) * Mp);
      } else if (mi->irgateMod == 3) {
        OR_ADD(F, S_(GEge), (mi->geltd) * Mp);
        OR_SUB(F, S_(GEgm), (mi->geltd) * Mp);
        OR_SUB(F, S_(GMge), (mi->geltd) * Mp);
        OR_ADD(F, S_(GMgm), (mi->geltd + mi->gcrg) * Mp);
        OR_ADD(F, S_(GMdp), (mi->gcrgd) * Mp);
        OR_ADD(F, S_(GMgp), (mi->gcrgg) * Mp);
        OR_ADD(F, S_(GMsp), (mi->gcrgs) * Mp);
        OR_ADD(F, S_(GMbp), (mi->gcrgb) * Mp);
        OR_SUB(F, S_(GPgm), (mi->gcrg) * Mp);
        OR_ADD(F, S_(GPgp), (-mi->gcrgg - mi->ggtg + mi->gIgtotg) * Mp);
        OR_ADD(F, S_(GPdp), (-mi->gcrgd - mi->ggtd + mi->gIgtotd) * Mp);
        OR_ADD(F, S_(GPsp), (-mi->gcrgs - mi->ggts + mi->gIgtots) * Mp);
        OR_ADD(F, S_(GPbp), (-mi->gcrgb - mi->ggtb + mi->gIgtotb) * Mp);
      } else {
        OR_ADD(F, S_(GPgp), (-mi->ggtg + mi->gIgtotg) * Mp);
        OR_ADD(F, S_(GPdp), (-mi->ggtd + mi->gIgtotd) * Mp);
        OR_ADD(F, S_(GPsp), (-mi->ggts + mi->gIgtots) * Mp);
        OR_ADD(F, S_(GPbp), (-mi->ggtb + mi->gIgtotb) * Mp);
      }
      if (M(rdsMod)) {
        OR_ADD(F, S_(Dgp), (mi->gdtotg) * Mp);
        OR_ADD(F, S_(Dsp), (mi->gdtots) * Mp);
        OR_ADD(F, S_(Dbp), (mi->gdtotb) * Mp);
        OR_ADD(F, S_(Sdp), (mi->gstotd) * Mp);
        OR_ADD(F, S_(Sgp), (mi->gstotg) * Mp);
        OR_ADD(F, S_(Sbp), (mi->gstotb) * Mp);
      }
      OR_ADD(F, S_(DPdp), (mi->gdpr + mi->gds + mi->gbd + T1 * mi->ddxpart_dVd - mi->gdtotd + mi->RevSum
                           + mi->gbdpdp + mi->dxpart * mi->ggtd - mi->gIdtotd) * Mp);
      OR_SUB(F, S_(DPd), (mi->gdpr + mi->gdtot) * Mp);
      OR_ADD(F, S_(DPgp), (mi->Gm - mi->gdtotg + mi->gbdpg - mi->gIdtotg + mi->dxpart * mi->ggtg
                           + T1 * mi->ddxpart_dVg) * Mp);
      OR_SUB(F, S_(DPsp), (mi->gds + mi->gdtots - mi->dxpart * mi->ggts + mi->gIdtots - T1 * mi->ddxpart_dVs
                           + mi->FwdSum - mi->gbdpsp) * Mp);
      OR_SUB(F, S_(DPbp), (mi->gjbd + mi->gdtotb - mi->Gmbs - mi->gbdpb + mi->gIdtotb - T1 * mi->ddxpart_dVb
                           - mi->dxpart * mi->ggtb) * Mp);
      OR_SUB(F, S_(Ddp), (mi->gdpr - mi->gdtotd) * Mp);
      OR_ADD(F, S_(Dd), (mi->gdpr + mi->gdtot) * Mp);
      OR_SUB(F, S_(SPdp), (mi->gds + mi->gstotd + mi->RevSum - mi->gbspdp - T1 * mi->dsxpart_dVd
                           - mi->sxpart * mi->ggtd + mi->gIstotd) * Mp);
      OR_ADD(F, S_(SPgp), (-mi->Gm - mi->gstotg + mi->gbspg + mi->sxpart * mi->ggtg + T1 * mi->dsxpart_dVg
                           - mi->gIstotg) * Mp);
      OR_ADD(F, S_(SPsp), (mi->gspr + mi->gds + mi->gbs + T1 * mi->dsxpart_dVs - mi->gstots + mi->FwdSum
                           + mi->gbspsp + mi->sxpart * mi->ggts - mi->gIstots) * Mp);
      OR_SUB(F, S_(SPs), (mi->gspr + mi->gstot) * Mp);
      OR_SUB(F, S_(SPbp), (mi->gjbs + mi->gstotb + mi->Gmbs - mi->gbspb - mi->sxpart * mi->ggtb
                           - T1 * mi->dsxpart_dVb + mi->gIstotb) * Mp);
      OR_SUB(F, S_(Ssp), (mi->gspr - mi->gstots) * Mp);
      OR_ADD(F, S_(Ss), (mi->gspr + mi->gstot) * Mp);
      OR_ADD(F, S_(BPdp), (-mi->gjbd + mi->gbbdp - mi->gIbtotd) * Mp);
      OR_ADD(F, S_(BPgp), (-mi->gbgs - mi->gIbtotg) * Mp);
      OR_ADD(F, S_(BPsp), (-mi->gjbs + mi->gbbsp - mi->gIbtots) * Mp);
      OR_ADD(F, S_(BPbp), (mi->gjbd + mi->gjbs - mi->gbbs - mi->gIbtotb) * Mp);
      OR_ADD(F, S_(DPdp), (mi->ggidld) * Mp);
      OR_ADD(F, S_(DPgp), (mi->ggidlg) * Mp);
      OR_SUB(F, S_(DPsp), ((mi->ggidlg + mi->ggidld + mi->ggidlb)) * Mp);
      OR_ADD(F, S_(DPbp), (mi->ggidlb) * Mp);
      OR_SUB(F, S_(BPdp), (mi->ggidld) * Mp);
      OR_SUB(F, S_(BPgp), (mi->ggidlg) * Mp);
      OR_ADD(F, S_(BPsp), ((mi->ggidlg + mi->ggidld + mi->ggidlb)) * Mp);
      OR_SUB(F, S_(BPbp), (mi->ggidlb) * Mp);
      OR_SUB(F, S_(SPdp), ((mi->ggisls + mi->ggislg + mi->ggislb)) * Mp);
      OR_ADD(F, S_(SPgp), (mi->ggislg) * Mp);
      OR_ADD(F, S_(SPsp), (mi->ggisls) * Mp);
      OR_ADD(F, S_(SPbp), (mi->ggislb) * Mp);
      OR_ADD(F, S_(BPdp), ((mi->ggislg + mi->ggisls + mi->ggislb)) * Mp);
      OR_SUB(F, S_(BPgp), (mi->ggislg) * Mp);
      OR_SUB(F, S_(BPsp), (mi->ggisls) * Mp);
      OR_SUB(F, S_(BPbp), (mi->ggislb) * Mp);
      if (mi->irbodyMod) {
        OR_ADD(F, S_(DPdb), (-mi->gbd) * Mp);
        OR_SUB(F, S_(SPsb), (mi->gbs) * Mp);
        OR_ADD(F, S_(DBdp), (-mi->gbd) * Mp);
        OR_ADD(F, S_(DBdb), (mi->gbd + mi->grbpd + mi->grbdb) * Mp);
        OR_SUB(F, S_(DBbp), (mi->grbpd) * Mp);
        OR_SUB(F, S_(DBb), (mi->grbdb) * Mp);
        OR_SUB(F, S_(BPdb), (mi->grbpd) * Mp);
        OR_SUB(F, S_(BPb), (mi->grbpb) * Mp);
        OR_SUB(F, S_(BPsb), (mi->grbps) * Mp);
        OR_ADD(F, S_(BPbp), (mi->grbpd + mi->grbps + mi->grbpb) * Mp);
        OR_ADD(F, S_(SBsp), (-mi->gbs) * Mp);
        OR_SUB(F, S_(SBbp), (mi->grbps) * Mp);
        OR_SUB(F, S_(SBb), (mi->grbsb) * Mp);
        OR_ADD(F, S_(SBsb), (mi->gbs + mi->grbps + mi->grbsb) * Mp);
        OR_SUB(F, S_(Bdb), (mi->grbdb) * Mp);
        OR_SUB(F, S_(Bbp), (mi->grbpb) * Mp);
        OR_SUB(F, S_(Bsb), (mi->grbsb) * Mp);
        OR_ADD(F, S_(Bb), (mi->grbsb + mi->grbdb + mi->grbpb) * Mp);
      }
      if (mi->itrnqsMod) {
        OR_ADD(F, S_(Qq), (mi->gqdef + mi->gtau) * Mp);
        OR_ADD(F, S_(Qgp), (mi->ggtg) * Mp);
        OR_ADD(F, S_(Qdp), (mi->ggtd) * Mp);
        OR_ADD(F, S_(Qsp), (mi->ggts) * Mp);
        OR_ADD(F, S_(Qbp), (mi->ggtb) * Mp);
        OR_ADD(F, S_(DPq), (mi->dxpart * mi->gtau) * Mp);
        OR_ADD(F, S_(SPq), (mi->sxpart * mi->gtau) * Mp);
        OR_SUB(F, S_(GPq), (mi->gtau) * Mp);
      }
      /* initial conditions (:16840-16887); the transient form assigns 1 to
       * the branch diagonal, which only this instance addresses */
      if (mi->icVBS_given) {
        if (ss->dcopFlag) {
          OR_ADD(F, S_(Bibs), 1.0);
          OR_ADD(F, S_(Sibs), -1.0);
          OR_ADD(F, S_(IBSb), 1.0);
          OR_ADD(F, S_(IBSs), -1.0);
        } else {
          OR_ADD(F, S_(IBSibs), 1.0);
        }
      }
      if (mi->icVDS_given) {
        if (ss->dcopFlag) {
          OR_ADD(F, S_(Dids), 1.0);
          OR_ADD(F, S_(Sids), -1.0);
          OR_ADD(F, S_(IDSd), 1.0);
          OR_ADD(F, S_(IDSs), -1.0);
        } else {
          OR_ADD(F, S_(IDSids), 1.0);
        }
      }
      if (mi->icVGS_given) {
        if (ss->dcopFlag) {
          OR_ADD(F, S_(GEigs), 1.0);
          OR_ADD(F, S_(Sigs), -1.0);
          OR_ADD(F, S_(IGSg), 1.0);
          OR_ADD(F, S_(IGSs), -1.0);
        } else {
          OR_ADD(F, S_(IGSigs), 1.0);
        }
      }
    }
    if (!(ss->dcopFlag) && ss->initTranFlag && ss->newtonIter == 0) {
      /* q is zero on the first transient iteration */
    } else if (Q) {
      if (mi->irgateMod == 3) {
        OR_ADD(Q, S_(GMgm), (+mi->CAPcgmgmb) * Mp);
        OR_ADD(Q, S_(GMdp), (mi->CAPcgmdb) * Mp);
        OR_ADD(Q, S_(GMsp), (mi->CAPcgmsb) * Mp);
        OR_ADD(Q, S_(GMbp), (mi->CAPcgmbb) * Mp);
        OR_ADD(Q, S_(DPgm), (mi->CAPcdgmb) * Mp);
        OR_ADD(Q, S_(SPgm), (mi->CAPcsgmb) * Mp);
        OR_ADD(Q, S_(BPgm), (mi->CAPcbgmb) * Mp);
      }
      OR_ADD(Q, S_(GPgp), (mi->CAPcggb) * Mp);
      OR_ADD(Q, S_(GPdp), (mi->CAPcgdb) * Mp);
      OR_ADD(Q, S_(GPsp), (mi->CAPcgsb) * Mp);
      OR_ADD(Q, S_(GPbp), (mi->CAPcgbb) * Mp);
      OR_ADD(Q, S_(DPdp), (mi->CAPcddb) * Mp);
      OR_ADD(Q, S_(DPgp), (+mi->CAPcdgb) * Mp);
      OR_SUB(Q, S_(DPsp), (-mi->CAPcdsb) * Mp);
      OR_SUB(Q, S_(DPbp), (-mi->CAPcdbb) * Mp);
      OR_SUB(Q, S_(SPdp), (-mi->CAPcsdb) * Mp);
      OR_ADD(Q, S_(SPgp), (mi->CAPcsgb) * Mp);
      OR_ADD(Q, S_(SPsp), (mi->CAPcssb) * Mp);
      OR_SUB(Q, S_(SPbp), (-mi->CAPcsbb) * Mp);
      OR_ADD(Q, S_(BPdp), (mi->CAPcbdb) * Mp);
      OR_ADD(Q, S_(BPgp), (mi->CAPcbgb) * Mp);
      OR_ADD(Q, S_(BPsp), (mi->CAPcbsb) * Mp);
      OR_ADD(Q, S_(BPbp), (mi->CAPcbbb) * Mp);
      if (mi->irbodyMod) {
        OR_ADD(Q, S_(DPdb), (mi->CAPcdbdb) * Mp);
        OR_SUB(Q, S_(SPsb), (-mi->CAPcsbsb) * Mp);
        OR_ADD(Q, S_(DBdp), (mi->CAPcdbdb) * Mp);
        OR_ADD(Q, S_(DBdb), (-mi->CAPcdbdb) * Mp);
        OR_ADD(Q, S_(SBsp), (mi->CAPcsbsb) * Mp);
        OR_ADD(Q, S_(SBsb), (-mi->CAPcsbsb) * Mp);
      }
      if (mi->itrnqsMod) {
        OR_ADD(Q, S_(Qgp), (-mi->CAPcqgb) * Mp);
        OR_ADD(Q, S_(Qdp), (-mi->CAPcqdb) * Mp);
        OR_ADD(Q, S_(Qsp), (-mi->CAPcqsb) * Mp);
        OR_ADD(Q, S_(Qbp), (-mi->CAPcqbb) * Mp);
      }
    }
#undef S_
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
    else if (!strcmp(name, "ua")) out[i] = p->ua;
    else if (!strcmp(name, "ub")) out[i] = p->ub;
    else if (!strcmp(name, "uc")) out[i] = p->uc;
    else if (!strcmp(name, "ud")) out[i] = p->ud;
    else if (!strcmp(name, "rds0")) out[i] = p->rds0;
    else if (!strcmp(name, "vfbzbfactor")) out[i] = p->vfbzbfactor;
    else if (!strcmp(name, "vtm")) out[i] = I->model->vtm;
    else if (!strcmp(name, "coxe")) out[i] = I->model->coxe;
    else if (!strcmp(name, "qsrc")) out[i] = I->qsrc;
    else if (!strcmp(name, "qdrn")) out[i] = I->qdrn;
    else if (!strcmp(name, "Igidl")) out[i] = I->Igidl;
    else if (!strcmp(name, "weffCJ")) out[i] = p->weffCJ;
    else if (!strcmp(name, "Igs")) out[i] = I->Igs;
    else if (!strcmp(name, "Igd")) out[i] = I->Igd;
    else if (!strcmp(name, "Igcs")) out[i] = I->Igcs;
    else if (!strcmp(name, "Igcd")) out[i] = I->Igcd;
    else if (!strcmp(name, "Igb")) out[i] = I->Igb;
    else if (!strcmp(name, "gstot")) out[i] = I->gstot;
    else if (!strcmp(name, "gdtot")) out[i] = I->gdtot;
    else if (!strcmp(name, "Igate")) out[i] = I->Igate;
    else if (!strcmp(name, "IgateMid")) out[i] = I->IgateMid;
    else if (!strcmp(name, "Idbb")) out[i] = I->Idbb;
    else if (!strcmp(name, "Isbb")) out[i] = I->Isbb;
    else if (!strcmp(name, "Ibpb")) out[i] = I->Ibpb;
    else if (!strcmp(name, "gcrg")) out[i] = I->gcrg;
    else if (!strcmp(name, "AechvbEdgeS")) out[i] = p->AechvbEdgeS;
    else if (!strcmp(name, "AechvbEdgeD")) out[i] = p->AechvbEdgeD;
    else if (!strcmp(name, "BechvbEdge")) out[i] = p->BechvbEdge;
    else if (!strcmp(name, "Aechvb")) out[i] = p->Aechvb;
    else if (!strcmp(name, "Bechvb")) out[i] = p->Bechvb;
    else if (!strcmp(name, "ToxRatio")) out[i] = p->ToxRatio;
    else if (!strcmp(name, "ToxRatioEdge")) out[i] = p->ToxRatioEdge;
    else if (!strcmp(name, "vfbsd")) out[i] = p->vfbsd;
    else if (!strcmp(name, "rs0")) out[i] = p->rs0;
    else if (!strcmp(name, "rswmin")) out[i] = p->rswmin;
    else if (!strcmp(name, "rd0")) out[i] = p->rd0;
    else if (!strcmp(name, "rdwmin")) out[i] = p->rdwmin;
    else if (!strcmp(name, "sourceConductance")) out[i] = I->sourceConductance;
    else if (!strcmp(name, "drainConductance")) out[i] = I->drainConductance;
    else if (!strcmp(name, "grgeltd")) out[i] = I->grgeltd;
    else if (!strcmp(name, "grbdb")) out[i] = I->grbdb;
    else if (!strcmp(name, "grbsb")) out[i] = I->grbsb;
    else if (!strcmp(name, "grbpb")) out[i] = I->grbpb;
    else if (!strcmp(name, "CoxWL")) out[i] = I->CoxWL;
    else if (!strcmp(name, "qbulk")) out[i] = I->qbulk;
    else if (!strcmp(name, "vfbcv")) out[i] = p->vfbcv;
    else if (!strcmp(name, "k1ox")) out[i] = p->k1ox;
    else if (!strcmp(name, "qgdo")) out[i] = I->qgdo;
    else if (!strcmp(name, "qgso")) out[i] = I->qgso;
    else if (!strcmp(name, "qgb")) out[i] = I->qgb;
    else if (!strcmp(name, "gtau")) out[i] = I->gtau;
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
