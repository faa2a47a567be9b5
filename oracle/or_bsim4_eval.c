/* or_bsim4_eval.c -- oracle restatement of BSIM4 Instance::updateIntermediateVars
 * (src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET_B4.C:7744-11919), test
 * infrastructure only.  Every branch of the reference is restated: mtrlMod,
 * capMod 0/1/2, igcMod 1/2, igbMod, rdsMod, rgateMod 1-3, rbodyMod,
 * trnqsMod/acnqsMod and the initial conditions. */
#include "or_bsim4.h"
#include "or_eval_libm.h"

#define B4_DEXP(A, B, C) do { \
    if ((A) > OR_EXP_THRESHOLD) { (B) = OR_MAX_EXP * (1.0 + (A) - OR_EXP_THRESHOLD); (C) = OR_MAX_EXP; } \
    else if ((A) < -OR_EXP_THRESHOLD) { (B) = OR_MIN_EXP; (C) = 0; } \
    else { (B) = exp(A); (C) = (B); } } while (0)

/* polyDepletion (N_DEV_MOSFET_B4.C:14252-14282) */
static void polyDepletion(double phi, double ngate, double epsgate, double coxe, double Vgs_arg,
                          double* Vgs_eff, double* dVgs_eff_dVg) {
  double T1, T2, T3, T4, T5, T6, T7, T8;
  if ((ngate > 1.0e18) && (ngate < 1.0e25) && (Vgs_arg > phi) && (epsgate != 0)) {
    T1 = 1.0e6 * OR_Q * epsgate * ngate / (coxe * coxe);
    T8 = Vgs_arg - phi;
    T4 = sqrt(1.0 + 2.0 * T8 / T1);
    T2 = 2.0 * T8 / (T4 + 1.0);
    T3 = 0.5 * T2 * T2 / T1;
    T7 = 1.12 - T3 - 0.05;
    T6 = sqrt(T7 * T7 + 0.224);
    T5 = 1.12 - 0.5 * (T7 + T6);
    *Vgs_eff = Vgs_arg - T5;
    *dVgs_eff_dVg = 1.0 - (0.5 - 0.5 / T4) * (1.0 + T7 / T6);
  } else {
    *Vgs_eff = Vgs_arg;
    *dVgs_eff_dVg = 1.0;
  }
}

static double sol(const double* x, int lid) { return lid >= 0 ? x[lid] : 0.0; }

int b4_update_intermediate(b4inst* I, const xa_solver_state* ss, const xa_device_options* opt,
                           const xa_vectors* v) {
  const b4model* m = I->model;
  const b4size* p = I->paramPtr;
  const double gmin = opt->gmin;
  const double nf = I->nf;
  double vgd_old, SourceSatCurrent, DrainSatCurrent, evbd, evbs, arg, sarg;
  double Vfbeff, dVfbeff_dVg, dVfbeff_dVb, V3, V4, MJD, MJSWD, MJSWGD, MJS, MJSWS, MJSWGS;
  double Ggidld, Ggidlg, Ggidlb, Ggislg, Ggislb, Ggisls, VgstNVt, ExpVgst;
  double czbd, czbdsw, czbdswg, czbs, czbssw, czbsswg;
  double Vgs_eff, dVgs_eff_dVg, Vth_NarrowW, Phis, dPhis_dVb, sqrtPhis, dsqrtPhis_dVb;
  double Vth, dVth_dVb, dVth_dVd, Vgst, Nvtms, Nvtmd, Vtm, Vtm0, n, dn_dVb, dn_dVd, voffcv;
  double noff, dnoff_dVd, dnoff_dVb, CoxWLcen, QovCox, LINK, V0, DeltaPhi, dDeltaPhi_dVg;
  double VgDP, dVgDP_dVg, Cox, Tox, Tcen, dTcen_dVg, dTcen_dVd, dTcen_dVb, Ccen;
  double Coxeff, dCoxeff_dVd, dCoxeff_dVg, dCoxeff_dVb, Denomi, dDenomi_dVg, dDenomi_dVd, dDenomi_dVb;
  double dueff_dVg, dueff_dVd, dueff_dVb, Esat, Vdsat, dEsatL_dVg, dEsatL_dVd, dEsatL_dVb;
  double dVdsat_dVg, dVdsat_dVb, dVdsat_dVd, Vasat, dVasat_dVg, dVasat_dVb, dVasat_dVd;
  double Va, dVa_dVd, dVa_dVg, dVa_dVb, Vbseff, dVbseff_dVb, VbseffCV, dVbseffCV_dVb;
  double T0, dT0_dVg, dT0_dVd, dT0_dVb, T1, dT1_dVg, dT1_dVd, dT1_dVb;
  double T2, dT2_dVg, dT2_dVd, dT2_dVb, T3, dT3_dVg, dT3_dVd, dT3_dVb;
  double T4, dT4_dVd, dT4_dVb, T5, dT5_dVg, dT5_dVd, dT5_dVb;
  double T6, dT6_dVg, dT6_dVd, dT6_dVb, T7, T8, dT8_dVg, dT8_dVd, dT8_dVb;
  double T9, dT9_dVg, dT9_dVd, dT9_dVb, T10, dT10_dVg, dT10_dVb, dT10_dVd;
  double T11, T12, T13, T14, tmp, dAbulk_dVb, Abulk0, dAbulk0_dVb;
  double Cclm, dCclm_dVg, dCclm_dVd, dCclm_dVb, FP, dFP_dVg, PvagTerm, dPvagTerm_dVg;
  double dPvagTerm_dVd, dPvagTerm_dVb, VADITS, dVADITS_dVg, dVADITS_dVd, Lpe_Vb;
  double dDITS_Sft_dVb, dDITS_Sft_dVd, VACLM, dVACLM_dVg, dVACLM_dVd, dVACLM_dVb;
  double VADIBL, dVADIBL_dVg, dVADIBL_dVd, dVADIBL_dVb, Xdep, dXdep_dVb;
  double lt1, dlt1_dVb, ltw, dltw_dVb, Delt_vth, dDelt_vth_dVb, Theta0, dTheta0_dVb;
  double TempRatio, tmp1, tmp2, tmp3, tmp4, DIBL_Sft, dDIBL_Sft_dVd, Lambda, dLambda_dVg, a1;
  double dVgsteff_dVg, dVgsteff_dVd, dVgsteff_dVb, dVdseff_dVg, dVdseff_dVd, dVdseff_dVb;
  double VdseffCV, dVdseffCV_dVg, dVdseffCV_dVd, dVdseffCV_dVb, diffVds, dAbulk_dVg;
  double beta, dbeta_dVg, dbeta_dVd, dbeta_dVb, gche, dgche_dVg, dgche_dVd, dgche_dVb;
  double fgche1, dfgche1_dVg, dfgche1_dVd, dfgche1_dVb, fgche2, dfgche2_dVg, dfgche2_dVd, dfgche2_dVb;
  double Idl, dIdl_dVg, dIdl_dVd, dIdl_dVb, Idsa, dIdsa_dVg, dIdsa_dVd, dIdsa_dVb, Ids, Gmb;
  double devbs_dvb, devbd_dvb, Isub, Gbd, Gbg, Gbb, Gds, VASCBE, dVASCBE_dVg, dVASCBE_dVd, dVASCBE_dVb;
  double CoxeffWovL, Rds, dRds_dVg, dRds_dVb, WVCox, WVCoxRds, Vgst2Vtm, VdsatCV;
  double Leff, Weff, dWeff_dVg, dWeff_dVb, AbulkCV, dAbulkCV_dVb;
  double Cgg1, Cgb1, Cgd1, Cbg1, Cbb1, Cbd1, Qac0, Qsub0, dQac0_dVg, dQac0_dVb;
  double dQsub0_dVg, dQsub0_dVd, dQsub0_dVb, Csg, Csd, Csb, Cbg, Cbd, Cbb, Cgg, Cgd, Cgb;
  double Nvtmrss, Nvtmrssws, Nvtmrsswgs, Nvtmrsd, Nvtmrsswd, Nvtmrsswgd;
  double vs, Fsevl, dvs_dVg, dvs_dVd, dvs_dVb, dFsevl_dVg, dFsevl_dVd, dFsevl_dVb;
  double vgdx, vgsx, epssub, toxe, epsrox, von_local, Vds, Vgs, Vbs, Vdb_, Gm_, Vdseff_;
  double vbd_old, vbs_old, vgs_old, vds_old, vges_old, vgms_old, vdes_old, vses_old;
  double vdbs_old, vsbs_old, vdbd_old, vged_old, vgmd_old, vsbd_old, vsbd;
  /* gate tunnelling, bias-dependent Rs/Rd, capMod 0/1 */
  double Vfb, Voxacc, dVoxacc_dVg, dVoxacc_dVb, Voxdepinv, dVoxdepinv_dVg, dVoxdepinv_dVd;
  double dVoxdepinv_dVb, VxNVt, ExpVxNVt, Vaux, dVaux_dVg, dVaux_dVd, dVaux_dVb, dT2_dVg_;
  double dT7_dVg, dT7_dVd, dT7_dVb, Igc, dIgc_dVg, dIgc_dVd, dIgc_dVb, Pigcd, dPigcd_dVg;
  double dPigcd_dVd, dPigcd_dVb, dIgcs_dVg, dIgcs_dVd, dIgcs_dVb, dIgcd_dVg, dIgcd_dVd, dIgcd_dVb;
  double dIgs_dVg, dIgs_dVs, dIgd_dVg, dIgd_dVd, Igbacc, dIgbacc_dVg, dIgbacc_dVb;
  double Igbinv, dIgbinv_dVg, dIgbinv_dVd, dIgbinv_dVb;
  double dT0_dvg, dT1_dvb, dT3_dvg, dT3_dvb, Rs, dRs_dvg, dRs_dvb, Rd, dRd_dvg, dRd_dvb;
  double dgstot_dvd, dgstot_dvg, dgstot_dvb, dgstot_dvs, dgdtot_dvd, dgdtot_dvg, dgdtot_dvb;
  double dgdtot_dvs, Arg1, One_Third_CoxWL, Two_Third_CoxWL, Alphaz, dAlphaz_dVg, dAlphaz_dVb;
  double dVgst_dVb, dVgst_dVg, dVdseffCV_dVg_, dVdseffCV_dVd_;
  int Check = 0, Check1 = 0, Check2 = 0;

  dT1_dVb = dT2_dVb = dT3_dVb = dT4_dVb = dT5_dVb = dT6_dVb = 0.0;
  dT0_dVb = 0.0;
  devbs_dvb = devbd_dvb = 0.0;
  I->ChargeComputationNeeded = (ss->tranopFlag || ss->acopFlag || ss->transientFlag) ? 1 : 0;
  I->limitedFlag = 0;

  /* node voltages (:7919-7974) */
  I->Vd = sol(v->nextSol, I->li_Drain);
  I->Vs = sol(v->nextSol, I->li_Source);
  I->Vb = sol(v->nextSol, I->li_Body);
  I->Vsp = sol(v->nextSol, I->li_SourcePrime);
  I->Vdp = sol(v->nextSol, I->li_DrainPrime);
  I->Vgp = sol(v->nextSol, I->li_GatePrime);
  I->Vbp = sol(v->nextSol, I->li_BodyPrime);
  I->Vge = sol(v->nextSol, I->li_GateExt);
  I->Vgm = (I->li_GateMid >= 0) ? sol(v->nextSol, I->li_GateMid) : 0.0;
  I->Vdb = sol(v->nextSol, I->li_DrainBody);
  I->Vsb = sol(v->nextSol, I->li_SourceBody);
  I->Qtotal = I->itrnqsMod ? sol(v->nextSol, I->li_Charge) : 0.0;
  I->Vddp = I->Vd - I->Vdp;
  I->Vssp = I->Vs - I->Vsp;
  I->Vdbb = I->Vdb - I->Vb;
  I->Vdbbp = I->Vdb - I->Vbp;
  I->Vsbb = I->Vsb - I->Vb;
  I->Vsbbp = I->Vsb - I->Vbp;
  I->Vbpb = I->Vbp - I->Vb;
  I->vds = m->dtype * (I->Vdp - I->Vsp);
  I->vgs = m->dtype * (I->Vgp - I->Vsp);
  I->vbs = m->dtype * (I->Vbp - I->Vsp);
  I->vges = m->dtype * (I->Vge - I->Vsp);
  I->vgms = m->dtype * (I->Vgm - I->Vsp);
  I->vdbs = m->dtype * (I->Vdb - I->Vsp);
  I->vsbs = m->dtype * (I->Vsb - I->Vsp);
  I->vses = m->dtype * (I->Vs - I->Vsp);
  I->vdes = m->dtype * (I->Vd - I->Vsp);
  I->qdef = m->dtype * (I->Qtotal);
  I->vbd = I->vbs - I->vds;
  I->vdbd = I->vdbs - I->vds;
  I->vgd = I->vgs - I->vds;
  I->vged = I->vges - I->vds;
  I->vgmd = I->vgms - I->vds;
  I->vgmb = I->vgms - I->vbs;
  I->vdbd = I->vdbs - I->vds;
  I->vbs_jct = (!I->irbodyMod) ? I->vbs : I->vsbs;
  I->vbd_jct = (!I->irbodyMod) ? I->vbd : I->vdbd;
  I->Vgegp = I->Vge - I->Vgp;
  I->Vgegm = I->Vge - I->Vgm;
  I->Vgmgp = I->Vgm - I->Vgp;
  I->origFlag = 1;
  I->vbd_orig = I->vbd;
  I->vbs_orig = I->vbs;
  I->vgs_orig = I->vgs;
  I->vds_orig = I->vds;
  I->vgd_orig = I->vgd;
  I->vges_orig = I->vges;
  I->vgms_orig = I->vgms;
  I->vdes_orig = I->vdes;
  I->vses_orig = I->vses;
  I->vdbs_orig = I->vdbs;
  I->vsbs_orig = I->vsbs;
  I->vdbd_orig = I->vdbd;
  I->vged_orig = I->vged;
  I->vgmd_orig = I->vgmd;
  I->vbs_jct_orig = I->vbs_jct;
  I->vbd_jct_orig = I->vbd_jct;
  I->vgmb_orig = I->vgmb;
  I->vgb_orig = I->vgb;        /* the previous call's vgb: vgb is updated below */
  I->Vgegp_orig = I->Vgegp;
  I->Vgegm_orig = I->Vgegm;
  I->Vgmgp_orig = I->Vgmgp;

  /* initial conditions (:8052-8097) */
  if (ss->initJctFlag && !I->OFF && opt->voltageLimiterFlag) {
    int setv = 1;
    if (ss->inputOPFlag) {
      const double* fl = v->flagSol;
      const int lids[11] = {I->li_Drain, I->li_GateExt, I->li_Source, I->li_Body, I->li_DrainPrime,
                            I->li_GatePrime, I->li_GateMid, I->li_SourcePrime, I->li_BodyPrime,
                            I->li_DrainBody, I->li_SourceBody};
      int k, any0 = (fl == NULL);
      for (k = 0; k < 11 && !any0; ++k) any0 = (lids[k] < 0) || fl[lids[k]] == 0;
      setv = any0;
    }
    if (setv) {
      I->vds = 0.1;
      I->vdes = 0.11;
      I->vses = -0.01;
      I->vgs = I->vges = I->vgms = m->dtype * I->vth0 + 0.1;
      I->origFlag = 0;
    }
    I->vbs = I->vdbs = I->vsbs = 0.0;
    I->vbd = I->vbs - I->vds;
    I->vdbd = I->vdbs - I->vds;
    I->vgd = I->vgs - I->vds;
    I->vged = I->vges - I->vds;
    I->vgmd = I->vgms - I->vds;
  } else if ((ss->initFixFlag || ss->initJctFlag) && I->OFF) {
    I->vds = I->vgs = I->vbs = I->vges = I->vgms = 0.0;
    I->vds = I->vsbs = I->vdes = I->vses = I->qdef = 0.0;
  }

  /* previous iteration's junction voltages (:8100-8158) */
  if (ss->newtonIter == 0) {
    I->newtonIterOld = 0;
    if (!ss->dcopFlag || (ss->locaEnabledFlag && ss->dcopFlag)) {
      const double* s = v->currSto;
      vbd_old = s[I->li_store_vbd];
      vbs_old = s[I->li_store_vbs];
      vgs_old = s[I->li_store_vgs];
      vds_old = s[I->li_store_vds];
      vges_old = s[I->li_store_vges];
      vgms_old = s[I->li_store_vgms];
      vdes_old = s[I->li_store_vdes];
      vses_old = s[I->li_store_vses];
      vdbs_old = s[I->li_store_vdbs];
      vsbs_old = s[I->li_store_vsbs];
      vdbd_old = s[I->li_store_vdbd];
      vged_old = s[I->li_store_vged];
      vgmd_old = s[I->li_store_vgmd];
      von_local = s[I->li_store_von];
    } else {
      vbd_old = I->vbd;
      vbs_old = I->vbs;
      vgs_old = I->vgs;
      vds_old = I->vds;
      vges_old = I->vges;
      vgms_old = I->vgms;
      vdes_old = I->vdes;
      vses_old = I->vses;
      vdbs_old = I->vdbs;
      vsbs_old = I->vsbs;
      vdbd_old = I->vdbd;
      vged_old = I->vged;
      vgmd_old = I->vgmd;
      von_local = 0.0;
    }
  } else {
    const double* s = v->nextSto;
    vbd_old = s[I->li_store_vbd];
    vbs_old = s[I->li_store_vbs];
    vgs_old = s[I->li_store_vgs];
    vds_old = s[I->li_store_vds];
    vges_old = s[I->li_store_vges];
    vgms_old = s[I->li_store_vgms];
    vdes_old = s[I->li_store_vdes];
    vses_old = s[I->li_store_vses];
    vdbs_old = s[I->li_store_vdbs];
    vsbs_old = s[I->li_store_vsbs];
    vdbd_old = s[I->li_store_vdbd];
    vged_old = s[I->li_store_vged];
    vgmd_old = s[I->li_store_vgmd];
    von_local = s[I->li_store_von];
  }
  vgd_old = vgs_old - vds_old;
  (void)vged_old; (void)vgmd_old; (void)vges_old; (void)vgms_old; (void)vdbs_old;
  (void)vsbs_old; (void)vdbd_old; (void)vbd_old;

  /* voltage limiting (:8165-8359) */
  if (opt->voltageLimiterFlag && !(ss->initFixFlag && I->OFF)) {
    if (ss->newtonIter >= 0 && !(ss->initJctFlag)) {
      if (vds_old >= 0.0) {
        I->vgs = or_fetlim(I->vgs, vgs_old, von_local);
        I->vds = I->vgs - I->vgd;
        I->vds = or_limvds(I->vds, vds_old);
        I->vgd = I->vgs - I->vds;
        if (I->irgateMod == 3) {
          I->vges = or_fetlim(I->vges, vges_old, von_local);
          I->vgms = or_fetlim(I->vgms, vgms_old, von_local);
        } else if ((I->irgateMod == 1) || (I->irgateMod == 2)) {
          I->vges = or_fetlim(I->vges, vges_old, von_local);
          I->vged = I->vges - I->vds;
        }
        if (M(rdsMod)) {
          I->vdes = or_limvds(I->vdes, vdes_old);
          I->vses = -or_limvds(-I->vses, -vses_old);
        }
      } else {
        I->vgd = or_fetlim(I->vgd, vgd_old, von_local);
        I->vds = I->vgs - I->vgd;
        I->vds = -or_limvds(-I->vds, -vds_old);
        I->vgs = I->vgd + I->vds;
        if (I->irgateMod == 3) {
          I->vged = or_fetlim(I->vged, vged_old, von_local);
          I->vges = I->vged + I->vds;
          I->vgmd = or_fetlim(I->vgmd, vgmd_old, von_local);
          I->vgms = I->vgmd + I->vds;
        }
        if ((I->irgateMod == 1) || (I->irgateMod == 2)) {
          I->vged = or_fetlim(I->vged, vged_old, von_local);
          I->vges = I->vged + I->vds;
        }
        if (M(rdsMod)) {
          I->vdes = -or_limvds(-I->vdes, -vdes_old);
          I->vses = or_limvds(I->vses, vses_old);
        }
      }
      if (I->vds >= 0.0) {
        I->vbs = or_pnjlim(I->vbs, vbs_old, B4_vt0, m->vcrit, &Check);
        I->vbd = I->vbs - I->vds;
        if (I->irbodyMod) {
          I->vdbs = or_pnjlim(I->vdbs, vdbs_old, B4_vt0, m->vcrit, &Check1);
          I->vdbd = I->vdbs - I->vds;
          I->vsbs = or_pnjlim(I->vsbs, vsbs_old, B4_vt0, m->vcrit, &Check2);
          if ((Check1 != 0) || (Check2 != 0)) Check = 1;
        }
      } else {
        I->vbd = or_pnjlim(I->vbd, vbd_old, B4_vt0, m->vcrit, &Check);
        I->vbs = I->vbd + I->vds;
        if (I->irbodyMod) {
          I->vdbd = or_pnjlim(I->vdbd, vdbd_old, B4_vt0, m->vcrit, &Check1);
          I->vdbs = I->vdbd + I->vds;
          vsbd_old = vsbs_old - vds_old;
          vsbd = I->vsbs - I->vds;
          vsbd = or_pnjlim(vsbd, vsbd_old, B4_vt0, m->vcrit, &Check2);
          I->vsbs = vsbd + I->vds;
          if ((Check1 != 0) || (Check2 != 0)) Check = 1;
        }
      }
    }
    if (Check == 1) I->limitedFlag = 1;
    {
      const double machprec = 2.220446049250313e-16;
      if (fabs(I->vbd_orig - I->vbd) > machprec || fabs(I->vbs_orig - I->vbs) > machprec ||
          fabs(I->vgs_orig - I->vgs) > machprec || fabs(I->vds_orig - I->vds) > machprec ||
          fabs(I->vgd_orig - I->vgd) > machprec || fabs(I->vges_orig - I->vges) > machprec ||
          fabs(I->vgms_orig - I->vgms) > machprec || fabs(I->vdes_orig - I->vdes) > machprec ||
          fabs(I->vses_orig - I->vses) > machprec || fabs(I->vdbs_orig - I->vdbs) > machprec ||
          fabs(I->vsbs_orig - I->vsbs) > machprec || fabs(I->vdbd_orig - I->vdbd) > machprec ||
          fabs(I->vbs_jct_orig - I->vbs_jct) > machprec || fabs(I->vbd_jct_orig - I->vbd_jct) > machprec ||
          fabs(I->vgmb_orig - I->vgmb) > machprec || fabs(I->vgb_orig - I->vgb) > machprec ||
          fabs(I->vged_orig - I->vged) > machprec || fabs(I->vgmd_orig - I->vgmd) > machprec)
        I->origFlag = 0;
    }
  }
  if (ss->newtonIter != 0 && ss->newtonIter != I->newtonIterOld) I->newtonIterOld = ss->newtonIter;
  I->vbd = I->vbs - I->vds;
  I->vgd = I->vgs - I->vds;
  I->vgb = I->vgs - I->vbs;
  I->vged = I->vges - I->vds;
  I->vgmd = I->vgms - I->vds;
  I->vgmb = I->vgms - I->vbs;
  I->vdbd = I->vdbs - I->vds;
  I->vbs_jct = (!I->irbodyMod) ? I->vbs : I->vsbs;
  I->vbd_jct = (!I->irbodyMod) ? I->vbd : I->vdbd;

  /* source/drain junction diodes (:8425-8769) */
  Nvtms = m->vtm * M(SjctEmissionCoeff);
  if ((I->Aseff <= 0.0) && (I->Pseff <= 0.0)) SourceSatCurrent = 1.0e-14;
  else SourceSatCurrent = I->Aseff * m->SjctTempSatCurDensity + I->Pseff * m->SjctSidewallTempSatCurDensity
                          + p->weffCJ * nf * m->SjctGateSidewallTempSatCurDensity;
  if (SourceSatCurrent <= 0.0) {
    I->gbs = gmin;
    I->cbs = I->gbs * I->vbs_jct;
  } else {
    switch ((int)M(dioMod)) {
      case 0:
        evbs = exp(I->vbs_jct / Nvtms);
        T1 = M(xjbvs) * exp(-(M(bvs) + I->vbs_jct) / Nvtms);
        I->gbs = SourceSatCurrent * (evbs + T1) / Nvtms + gmin;
        I->cbs = SourceSatCurrent * (evbs + I->XExpBVS - T1 - 1.0) + gmin * I->vbs_jct;
        break;
      case 1:
        T2 = I->vbs_jct / Nvtms;
        if (T2 < -OR_EXP_THRESHOLD) {
          I->gbs = gmin;
          I->cbs = SourceSatCurrent * (OR_MIN_EXP - 1.0) + gmin * I->vbs_jct;
        } else if (I->vbs_jct <= I->vjsmFwd) {
          evbs = exp(T2);
          I->gbs = SourceSatCurrent * evbs / Nvtms + gmin;
          I->cbs = SourceSatCurrent * (evbs - 1.0) + gmin * I->vbs_jct;
        } else {
          T0 = I->IVjsmFwd / Nvtms;
          I->gbs = T0 + gmin;
          I->cbs = I->IVjsmFwd - SourceSatCurrent + T0 * (I->vbs_jct - I->vjsmFwd) + gmin * I->vbs_jct;
        }
        break;
      case 2:
        if (I->vbs_jct < I->vjsmRev) {
          T0 = I->vbs_jct / Nvtms;
          if (T0 < -OR_EXP_THRESHOLD) {
            evbs = OR_MIN_EXP;
            devbs_dvb = 0.0;
          } else {
            evbs = exp(T0);
            devbs_dvb = evbs / Nvtms;
          }
          T1 = evbs - 1.0;
          T2 = I->IVjsmRev + I->SslpRev * (I->vbs_jct - I->vjsmRev);
          I->gbs = devbs_dvb * T2 + T1 * I->SslpRev + gmin;
          I->cbs = T1 * T2 + gmin * I->vbs_jct;
        } else if (I->vbs_jct <= I->vjsmFwd) {
          T0 = I->vbs_jct / Nvtms;
          if (T0 < -OR_EXP_THRESHOLD) {
            evbs = OR_MIN_EXP;
            devbs_dvb = 0.0;
          } else {
            evbs = exp(T0);
            devbs_dvb = evbs / Nvtms;
          }
          T1 = (M(bvs) + I->vbs_jct) / Nvtms;
          if (T1 > OR_EXP_THRESHOLD) {
            T2 = OR_MIN_EXP;
            T3 = 0.0;
          } else {
            T2 = exp(-T1);
            T3 = -T2 / Nvtms;
          }
          I->gbs = SourceSatCurrent * (devbs_dvb - M(xjbvs) * T3) + gmin;
          I->cbs = SourceSatCurrent * (evbs + I->XExpBVS - 1.0 - M(xjbvs) * T2) + gmin * I->vbs_jct;
        } else {
          I->gbs = I->SslpFwd + gmin;
          I->cbs = I->IVjsmFwd + I->SslpFwd * (I->vbs_jct - I->vjsmFwd) + gmin * I->vbs_jct;
        }
        break;
      default: break;
    }
  }
  Nvtmd = m->vtm * M(DjctEmissionCoeff);
  if ((I->Adeff <= 0.0) && (I->Pdeff <= 0.0)) DrainSatCurrent = 1.0e-14;
  else DrainSatCurrent = I->Adeff * m->DjctTempSatCurDensity + I->Pdeff * m->DjctSidewallTempSatCurDensity
                         + p->weffCJ * nf * m->DjctGateSidewallTempSatCurDensity;
  if (DrainSatCurrent <= 0.0) {
    I->gbd = gmin;
    I->cbd = I->gbd * I->vbd_jct;
  } else {
    switch ((int)M(dioMod)) {
      case 0:
        evbd = exp(I->vbd_jct / Nvtmd);
        T1 = M(xjbvd) * exp(-(M(bvd) + I->vbd_jct) / Nvtmd);
        I->gbd = DrainSatCurrent * (evbd + T1) / Nvtmd + gmin;
        I->cbd = DrainSatCurrent * (evbd + I->XExpBVD - T1 - 1.0) + gmin * I->vbd_jct;
        break;
      case 1:
        T2 = I->vbd_jct / Nvtmd;
        if (T2 < -OR_EXP_THRESHOLD) {
          I->gbd = gmin;
          I->cbd = DrainSatCurrent * (OR_MIN_EXP - 1.0) + gmin * I->vbd_jct;
        } else if (I->vbd_jct <= I->vjdmFwd) {
          evbd = exp(T2);
          I->gbd = DrainSatCurrent * evbd / Nvtmd + gmin;
          I->cbd = DrainSatCurrent * (evbd - 1.0) + gmin * I->vbd_jct;
        } else {
          T0 = I->IVjdmFwd / Nvtmd;
          I->gbd = T0 + gmin;
          I->cbd = I->IVjdmFwd - DrainSatCurrent + T0 * (I->vbd_jct - I->vjdmFwd) + gmin * I->vbd_jct;
        }
        break;
      case 2:
        if (I->vbd_jct < I->vjdmRev) {
          T0 = I->vbd_jct / Nvtmd;
          if (T0 < -OR_EXP_THRESHOLD) {
            evbd = OR_MIN_EXP;
            devbd_dvb = 0.0;
          } else {
            evbd = exp(T0);
            devbd_dvb = evbd / Nvtmd;
          }
          T1 = evbd - 1.0;
          T2 = I->IVjdmRev + I->DslpRev * (I->vbd_jct - I->vjdmRev);
          I->gbd = devbd_dvb * T2 + T1 * I->DslpRev + gmin;
          I->cbd = T1 * T2 + gmin * I->vbd_jct;
        } else if (I->vbd_jct <= I->vjdmFwd) {
          T0 = I->vbd_jct / Nvtmd;
          if (T0 < -OR_EXP_THRESHOLD) {
            evbd = OR_MIN_EXP;
            devbd_dvb = 0.0;
          } else {
            evbd = exp(T0);
            devbd_dvb = evbd / Nvtmd;
          }
          T1 = (M(bvd) + I->vbd_jct) / Nvtmd;
          if (T1 > OR_EXP_THRESHOLD) {
            T2 = OR_MIN_EXP;
            T3 = 0.0;
          } else {
            T2 = exp(-T1);
            T3 = -T2 / Nvtmd;
          }
          I->gbd = DrainSatCurrent * (devbd_dvb - M(xjbvd) * T3) + gmin;
          I->cbd = DrainSatCurrent * (evbd + I->XExpBVD - 1.0 - M(xjbvd) * T2) + gmin * I->vbd_jct;
        } else {
          I->gbd = I->DslpFwd + gmin;
          I->cbd = I->IVjdmFwd + I->DslpFwd * (I->vbd_jct - I->vjdmFwd) + gmin * I->vbd_jct;
        }
        break;
      default: break;
    }
  }

  /* trap-assisted tunneling and recombination (:8655-8769) */
  Nvtmrssws = m->vtm0 * m->njtsswstemp;
  Nvtmrsswgs = m->vtm0 * m->njtsswgstemp;
  Nvtmrss = m->vtm0 * m->njtsstemp;
  Nvtmrsswd = m->vtm0 * m->njtsswdtemp;
  Nvtmrsswgd = m->vtm0 * m->njtsswgdtemp;
  Nvtmrsd = m->vtm0 * m->njtsdtemp;
  if ((M(vtss) - I->vbs_jct) < (M(vtss) * 1e-3)) {
    T9 = 1.0e3;
    T0 = -I->vbs_jct / Nvtmrss * T9;
    B4_DEXP(T0, T1, T10);
    dT1_dVb = T10 / Nvtmrss * T9;
  } else {
    T9 = 1.0 / (M(vtss) - I->vbs_jct);
    T0 = -I->vbs_jct / Nvtmrss * M(vtss) * T9;
    dT0_dVb = M(vtss) / Nvtmrss * (T9 + I->vbs_jct * T9 * T9);
    B4_DEXP(T0, T1, T10);
    dT1_dVb = T10 * dT0_dVb;
  }
  if ((M(vtsd) - I->vbd_jct) < (M(vtsd) * 1e-3)) {
    T9 = 1.0e3;
    T0 = -I->vbd_jct / Nvtmrsd * T9;
    B4_DEXP(T0, T2, T10);
    dT2_dVb = T10 / Nvtmrsd * T9;
  } else {
    T9 = 1.0 / (M(vtsd) - I->vbd_jct);
    T0 = -I->vbd_jct / Nvtmrsd * M(vtsd) * T9;
    dT0_dVb = M(vtsd) / Nvtmrsd * (T9 + I->vbd_jct * T9 * T9);
    B4_DEXP(T0, T2, T10);
    dT2_dVb = T10 * dT0_dVb;
  }
  if ((M(vtssws) - I->vbs_jct) < (M(vtssws) * 1e-3)) {
    T9 = 1.0e3;
    T0 = -I->vbs_jct / Nvtmrssws * T9;
    B4_DEXP(T0, T3, T10);
    dT3_dVb = T10 / Nvtmrssws * T9;
  } else {
    T9 = 1.0 / (M(vtssws) - I->vbs_jct);
    T0 = -I->vbs_jct / Nvtmrssws * M(vtssws) * T9;
    dT0_dVb = M(vtssws) / Nvtmrssws * (T9 + I->vbs_jct * T9 * T9);
    B4_DEXP(T0, T3, T10);
    dT3_dVb = T10 * dT0_dVb;
  }
  if ((M(vtsswd) - I->vbd_jct) < (M(vtsswd) * 1e-3)) {
    T9 = 1.0e3;
    T0 = -I->vbd_jct / Nvtmrsswd * T9;
    B4_DEXP(T0, T4, T10);
    dT4_dVb = T10 / Nvtmrsswd * T9;
  } else {
    T9 = 1.0 / (M(vtsswd) - I->vbd_jct);
    T0 = -I->vbd_jct / Nvtmrsswd * M(vtsswd) * T9;
    dT0_dVb = M(vtsswd) / Nvtmrsswd * (T9 + I->vbd_jct * T9 * T9);
    B4_DEXP(T0, T4, T10);
    dT4_dVb = T10 * dT0_dVb;
  }
  if ((M(vtsswgs) - I->vbs_jct) < (M(vtsswgs) * 1e-3)) {
    T9 = 1.0e3;
    T0 = -I->vbs_jct / Nvtmrsswgs * T9;
    B4_DEXP(T0, T5, T10);
    dT5_dVb = T10 / Nvtmrsswgs * T9;
  } else {
    T9 = 1.0 / (M(vtsswgs) - I->vbs_jct);
    T0 = -I->vbs_jct / Nvtmrsswgs * M(vtsswgs) * T9;
    dT0_dVb = M(vtsswgs) / Nvtmrsswgs * (T9 + I->vbs_jct * T9 * T9);
    B4_DEXP(T0, T5, T10);
    dT5_dVb = T10 * dT0_dVb;
  }
  if ((M(vtsswgd) - I->vbd_jct) < (M(vtsswgd) * 1e-3)) {
    T9 = 1.0e3;
    T0 = -I->vbd_jct / Nvtmrsswgd * T9;
    B4_DEXP(T0, T6, T10);
    dT6_dVb = T10 / Nvtmrsswgd * T9;
  } else {
    T9 = 1.0 / (M(vtsswgd) - I->vbd_jct);
    T0 = -I->vbd_jct / Nvtmrsswgd * M(vtsswgd) * T9;
    dT0_dVb = M(vtsswgd) / Nvtmrsswgd * (T9 + I->vbd_jct * T9 * T9);
    B4_DEXP(T0, T6, T10);
    dT6_dVb = T10 * dT0_dVb;
  }
  I->gbs += I->SjctTempRevSatCur * dT1_dVb + I->SswTempRevSatCur * dT3_dVb
            + I->SswgTempRevSatCur * dT5_dVb;
  I->cbs -= I->SjctTempRevSatCur * (T1 - 1.0) + I->SswTempRevSatCur * (T3 - 1.0)
            + I->SswgTempRevSatCur * (T5 - 1.0);
  I->gbd += I->DjctTempRevSatCur * dT2_dVb + I->DswTempRevSatCur * dT4_dVb
            + I->DswgTempRevSatCur * dT6_dVb;
  I->cbd -= I->DjctTempRevSatCur * (T2 - 1.0) + I->DswTempRevSatCur * (T4 - 1.0)
            + I->DswgTempRevSatCur * (T6 - 1.0);

  /* mode (:8773-8788) */
  if (I->vds >= 0.0) {
    I->mode = 1;
    Vds = I->vds;
    Vgs = I->vgs;
    Vbs = I->vbs;
    Vdb_ = I->vds - I->vbs;
  } else {
    I->mode = -1;
    Vds = -I->vds;
    Vgs = I->vgd;
    Vbs = I->vbd;
    Vdb_ = -I->vbs;
  }
  (void)Vdb_;
  if (M(mtrlMod)) {
    epsrox = 3.9;
    toxe = M(eot);
    epssub = B4_EPS0 * M(epsrsub);
  } else {
    epsrox = M(epsrox);
    toxe = M(toxe);
    epssub = B4_EPSSI;
  }
  /* MOSFET homotopy (:8818-8838), one gain-scale block, no staggering */
  if (ss->artParameterFlag) {
    const double alpha = ss->gainScale;
    const double vgstConst = 4.5, vdsScaleMin = 0.3;
    Vds = Vds * (ss->nltermScale * (1.0 - vdsScaleMin) + vdsScaleMin);
    Vgs = (alpha) * Vgs + (1.0 - alpha) * vgstConst;
  }

  /* Vbseff (:8849-8868) */
  T0 = Vbs - I->vbsc - 0.001;
  T1 = sqrt(T0 * T0 - 0.004 * I->vbsc);
  if (T0 >= 0.0) {
    Vbseff = I->vbsc + 0.5 * (T0 + T1);
    dVbseff_dVb = 0.5 * (1.0 + T0 / T1);
  } else {
    T2 = -0.002 / (T1 - T0);
    Vbseff = I->vbsc * (1.0 + T2);
    dVbseff_dVb = T2 * I->vbsc / T1;
  }
  T9 = 0.95 * p->phi;
  T0 = T9 - Vbseff - 0.001;
  T1 = sqrt(T0 * T0 + 0.004 * T9);
  Vbseff = T9 - 0.5 * (T0 + T1);
  dVbseff_dVb *= 0.5 * (1.0 + T0 / T1);
  Phis = p->phi - Vbseff;
  dPhis_dVb = -1.0;
  sqrtPhis = sqrt(Phis);
  dsqrtPhis_dVb = -0.5 / sqrtPhis;
  Xdep = p->Xdep0 * sqrtPhis / p->sqrtPhi;
  dXdep_dVb = (p->Xdep0 / p->sqrtPhi) * dsqrtPhis_dVb;
  Leff = p->leff;
  Vtm = m->vtm;
  Vtm0 = m->vtm0;

  /* threshold voltage (:8883-9052) */
  T3 = sqrt(Xdep);
  V0 = p->vbi - p->phi;
  T0 = p->dvt2 * Vbseff;
  if (T0 >= -0.5) {
    T1 = 1.0 + T0;
    T2 = p->dvt2;
  } else {
    T4 = 1.0 / (3.0 + 8.0 * T0);
    T1 = (1.0 + 3.0 * T0) * T4;
    T2 = p->dvt2 * T4 * T4;
  }
  lt1 = m->factor1 * T3 * T1;
  dlt1_dVb = m->factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
  T0 = p->dvt2w * Vbseff;
  if (T0 >= -0.5) {
    T1 = 1.0 + T0;
    T2 = p->dvt2w;
  } else {
    T4 = 1.0 / (3.0 + 8.0 * T0);
    T1 = (1.0 + 3.0 * T0) * T4;
    T2 = p->dvt2w * T4 * T4;
  }
  ltw = m->factor1 * T3 * T1;
  dltw_dVb = m->factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
  T0 = p->dvt1 * Leff / lt1;
  if (T0 < OR_EXP_THRESHOLD) {
    T1 = exp(T0);
    T2 = T1 - 1.0;
    T3 = T2 * T2;
    T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
    Theta0 = T1 / T4;
    dT1_dVb = -T0 * T1 * dlt1_dVb / lt1;
    dTheta0_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + OR_MIN_EXP)) / T4 / T4;
  } else {
    Theta0 = 1.0 / (OR_MAX_EXP - 2.0);
    dTheta0_dVb = 0.0;
  }
  I->thetavth = p->dvt0 * Theta0;
  Delt_vth = I->thetavth * V0;
  dDelt_vth_dVb = p->dvt0 * dTheta0_dVb * V0;
  T0 = p->dvt1w * p->weff * Leff / ltw;
  if (T0 < OR_EXP_THRESHOLD) {
    T1 = exp(T0);
    T2 = T1 - 1.0;
    T3 = T2 * T2;
    T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
    T5 = T1 / T4;
    dT1_dVb = -T0 * T1 * dltw_dVb / ltw;
    dT5_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + OR_MIN_EXP)) / T4 / T4;
  } else {
    T5 = 1.0 / (OR_MAX_EXP - 2.0);
    dT5_dVb = 0.0;
  }
  T0 = p->dvt0w * T5;
  T2 = T0 * V0;
  dT2_dVb = p->dvt0w * dT5_dVb * V0;
  TempRatio = I->temp / M(tnom) - 1.0;
  T0 = sqrt(1.0 + p->lpe0 / Leff);
  T1 = p->k1ox * (T0 - 1.0) * p->sqrtPhi + (p->kt1 + p->kt1l / Leff + p->kt2 * Vbseff) * TempRatio;
  Vth_NarrowW = toxe * p->phi / (p->weff + p->w0);
  T3 = I->eta0 + p->etab * Vbseff;
  if (T3 < 1.0e-4) {
    T9 = 1.0 / (3.0 - 2.0e4 * T3);
    T3 = (2.0e-4 - T3) * T9;
    T4 = T9 * T9;
  } else {
    T4 = 1.0;
  }
  dDIBL_Sft_dVd = T3 * p->theta0vb0;
  DIBL_Sft = dDIBL_Sft_dVd * Vds;
  Lpe_Vb = sqrt(1.0 + p->lpeb / Leff);
  Vth = m->dtype * I->vth0 + (p->k1ox * sqrtPhis - p->k1 * p->sqrtPhi) * Lpe_Vb
        - I->k2ox * Vbseff - Delt_vth - T2 + (p->k3 + p->k3b * Vbseff) * Vth_NarrowW + T1 - DIBL_Sft;
  dVth_dVb = Lpe_Vb * p->k1ox * dsqrtPhis_dVb - I->k2ox - dDelt_vth_dVb - dT2_dVb
             + p->k3b * Vth_NarrowW - p->etab * Vds * p->theta0vb0 * T4 + p->kt2 * TempRatio;
  dVth_dVd = -dDIBL_Sft_dVd;
  tmp1 = epssub / Xdep;
  I->nstar = m->vtm / OR_Q * (m->coxe + tmp1 + p->cit);
  tmp2 = p->nfactor * tmp1;
  tmp3 = p->cdsc + p->cdscb * Vbseff + p->cdscd * Vds;
  tmp4 = (tmp2 + tmp3 * Theta0 + p->cit) / m->coxe;
  if (tmp4 >= -0.5) {
    n = 1.0 + tmp4;
    dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + p->cdscb * Theta0) / m->coxe;
    dn_dVd = p->cdscd * Theta0 / m->coxe;
  } else {
    T0 = 1.0 / (3.0 + 8.0 * tmp4);
    n = (1.0 + 3.0 * tmp4) * T0;
    T0 *= T0;
    dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + p->cdscb * Theta0) / m->coxe * T0;
    dn_dVd = p->cdscd * Theta0 / m->coxe * T0;
  }
  if (p->dvtp0 > 0.0) {
    T0 = -p->dvtp1 * Vds;
    if (T0 < -OR_EXP_THRESHOLD) {
      T2 = OR_MIN_EXP;
      dT2_dVd = 0.0;
    } else {
      T2 = exp(T0);
      dT2_dVd = -p->dvtp1 * T2;
    }
    T3 = Leff + p->dvtp0 * (1.0 + T2);
    dT3_dVd = p->dvtp0 * dT2_dVd;
    if (M(tempMod) < 2) {
      T4 = Vtm * log(Leff / T3);
      dT4_dVd = -Vtm * dT3_dVd / T3;
    } else {
      T4 = m->vtm0 * log(Leff / T3);
      dT4_dVd = -m->vtm0 * dT3_dVd / T3;
    }
    dDITS_Sft_dVd = dn_dVd * T4 + n * dT4_dVd;
    dDITS_Sft_dVb = T4 * dn_dVb;
    Vth -= n * T4;
    dVth_dVd -= dDITS_Sft_dVd;
    dVth_dVb -= dDITS_Sft_dVb;
  }
  I->von = Vth;

  /* poly gate depletion (:9055-9080) */
  T0 = I->vfb + p->phi;
  if (M(mtrlMod) == 0) T1 = B4_EPSSI;
  else T1 = M(epsrgate) * B4_EPS0;
  polyDepletion(T0, p->ngate, T1, m->coxe, I->vgs, &I->vgs_eff, &I->dvgs_eff_dvg);
  polyDepletion(T0, p->ngate, T1, m->coxe, I->vgd, &I->vgd_eff, &I->dvgd_eff_dvg);
  if (I->mode > 0) {
    Vgs_eff = I->vgs_eff;
    dVgs_eff_dVg = I->dvgs_eff_dvg;
  } else {
    Vgs_eff = I->vgd_eff;
    dVgs_eff_dVg = I->dvgd_eff_dvg;
  }
  Vgst = Vgs_eff - Vth;

  /* effective Vgst (:9085-9150) */
  T0 = n * Vtm;
  T1 = p->mstar * Vgst;
  T2 = T1 / T0;
  if (T2 > OR_EXP_THRESHOLD) {
    T10 = T1;
    dT10_dVg = p->mstar * dVgs_eff_dVg;
    dT10_dVd = -dVth_dVd * p->mstar;
    dT10_dVb = -dVth_dVb * p->mstar;
  } else if (T2 < -OR_EXP_THRESHOLD) {
    T10 = Vtm * log(1.0 + OR_MIN_EXP);
    dT10_dVg = 0.0;
    dT10_dVd = T10 * dn_dVd;
    dT10_dVb = T10 * dn_dVb;
    T10 *= n;
  } else {
    ExpVgst = exp(T2);
    T3 = Vtm * log(1.0 + ExpVgst);
    T10 = n * T3;
    dT10_dVg = p->mstar * ExpVgst / (1.0 + ExpVgst);
    dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
    dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
    dT10_dVg *= dVgs_eff_dVg;
  }
  T1 = p->voffcbn - (1.0 - p->mstar) * Vgst;
  T2 = T1 / T0;
  if (T2 < -OR_EXP_THRESHOLD) {
    T3 = m->coxe * OR_MIN_EXP / p->cdep0;
    T9 = p->mstar + T3 * n;
    dT9_dVg = 0.0;
    dT9_dVd = dn_dVd * T3;
    dT9_dVb = dn_dVb * T3;
  } else if (T2 > OR_EXP_THRESHOLD) {
    T3 = m->coxe * OR_MAX_EXP / p->cdep0;
    T9 = p->mstar + T3 * n;
    dT9_dVg = 0.0;
    dT9_dVd = dn_dVd * T3;
    dT9_dVb = dn_dVb * T3;
  } else {
    ExpVgst = exp(T2);
    T3 = m->coxe / p->cdep0;
    T4 = T3 * ExpVgst;
    T5 = T1 * T4 / T0;
    T9 = p->mstar + n * T4;
    dT9_dVg = T3 * (p->mstar - 1.0) * ExpVgst / Vtm;
    dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
    dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
    dT9_dVg *= dVgs_eff_dVg;
  }
  I->Vgsteff = T10 / T9;
  T11 = T9 * T9;
  dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
  dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
  dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;

  /* effective channel geometry and series resistance (:9153-9198) */
  T9 = sqrtPhis - p->sqrtPhi;
  Weff = p->weff - 2.0 * (p->dwg * I->Vgsteff + p->dwb * T9);
  dWeff_dVg = -2.0 * p->dwg;
  dWeff_dVb = -2.0 * p->dwb * dsqrtPhis_dVb;
  if (Weff < 2.0e-8) {
    T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
    Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
    T0 *= T0 * 4.0e-16;
    dWeff_dVg *= T0;
    dWeff_dVb *= T0;
  }
  if (M(rdsMod) == 1) {
    Rds = dRds_dVg = dRds_dVb = 0.0;
  } else {
    T0 = 1.0 + p->prwg * I->Vgsteff;
    dT0_dVg = -p->prwg / T0 / T0;
    T1 = p->prwb * T9;
    dT1_dVb = p->prwb * dsqrtPhis_dVb;
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    dT3_dVg = 1.0 + T2 / (T3 - T2);
    dT3_dVb = dT3_dVg * dT1_dVb;
    dT3_dVg *= dT0_dVg;
    T4 = p->rds0 * 0.5;
    Rds = p->rdswmin + T3 * T4;
    dRds_dVg = T4 * dT3_dVg;
    dRds_dVb = T4 * dT3_dVb;
    if (Rds > 0.0) I->grdsw = 1.0 / Rds;
    else I->grdsw = 0.0;
  }

  /* bulk charge effect (:9201-9258) */
  T9 = 0.5 * p->k1ox * Lpe_Vb / sqrtPhis;
  T1 = T9 + I->k2ox - p->k3b * Vth_NarrowW;
  dT1_dVb = -T9 / sqrtPhis * dsqrtPhis_dVb;
  T9 = sqrt(p->xj * Xdep);
  tmp1 = Leff + 2.0 * T9;
  T5 = Leff / tmp1;
  tmp2 = p->a0 * T5;
  tmp3 = p->weff + p->b1;
  tmp4 = p->b0 / tmp3;
  T2 = tmp2 + tmp4;
  dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
  T6 = T5 * T5;
  T7 = T5 * T6;
  Abulk0 = 1.0 + T1 * T2;
  dAbulk0_dVb = T1 * tmp2 * dT2_dVb + T2 * dT1_dVb;
  T8 = p->ags * p->a0 * T7;
  dAbulk_dVg = -T1 * T8;
  I->Abulk = Abulk0 + dAbulk_dVg * I->Vgsteff;
  dAbulk_dVb = dAbulk0_dVb - T8 * I->Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb);
  if (Abulk0 < 0.1) {
    T9 = 1.0 / (3.0 - 20.0 * Abulk0);
    Abulk0 = (0.2 - Abulk0) * T9;
    dAbulk0_dVb *= T9 * T9;
  }
  if (I->Abulk < 0.1) {
    T9 = 1.0 / (3.0 - 20.0 * I->Abulk);
    I->Abulk = (0.2 - I->Abulk) * T9;
    T10 = T9 * T9;
    dAbulk_dVb *= T10;
    dAbulk_dVg *= T10;
  }
  T2 = p->keta * Vbseff;
  if (T2 >= -0.9) {
    T0 = 1.0 / (1.0 + T2);
    dT0_dVb = -p->keta * T0 * T0;
  } else {
    T1 = 1.0 / (0.8 + T2);
    T0 = (17.0 + 20.0 * T2) * T1;
    dT0_dVb = -p->keta * T1 * T1;
  }
  dAbulk_dVg *= T0;
  dAbulk_dVb = dAbulk_dVb * T0 + I->Abulk * dT0_dVb;
  dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVb;
  I->Abulk *= T0;
  Abulk0 *= T0;

  /* mobility (:9261-9341) */
  if (M(mtrlMod)) T14 = 2.0 * m->dtype * (M(phig) - M(easub) - 0.5 * m->Eg0 + 0.45);
  else T14 = 0.0;
  if (M(mobMod) == 0) {
    T0 = I->Vgsteff + Vth + Vth - T14;
    T2 = p->ua + p->uc * Vbseff;
    T3 = T0 / toxe;
    T12 = sqrt(Vth * Vth + 0.0001);
    T9 = 1.0 / (I->Vgsteff + 2 * T12);
    T10 = T9 * toxe;
    T8 = p->ud * T10 * T10 * Vth;
    T6 = T8 * Vth;
    T5 = T3 * (T2 + p->ub * T3) + T6;
    T7 = -2.0 * T6 * T9;
    T11 = T7 * Vth / T12;
    dDenomi_dVg = (T2 + 2.0 * p->ub * T3) / toxe;
    T13 = 2.0 * (dDenomi_dVg + T11 + T8);
    dDenomi_dVd = T13 * dVth_dVd;
    dDenomi_dVb = T13 * dVth_dVb + p->uc * T3;
    dDenomi_dVg += T7;
  } else if (M(mobMod) == 1) {
    T0 = I->Vgsteff + Vth + Vth - T14;
    T2 = 1.0 + p->uc * Vbseff;
    T3 = T0 / toxe;
    T4 = T3 * (p->ua + p->ub * T3);
    T12 = sqrt(Vth * Vth + 0.0001);
    T9 = 1.0 / (I->Vgsteff + 2 * T12);
    T10 = T9 * toxe;
    T8 = p->ud * T10 * T10 * Vth;
    T6 = T8 * Vth;
    T5 = T4 * T2 + T6;
    T7 = -2.0 * T6 * T9;
    T11 = T7 * Vth / T12;
    dDenomi_dVg = (p->ua + 2.0 * p->ub * T3) * T2 / toxe;
    T13 = 2.0 * (dDenomi_dVg + T11 + T8);
    dDenomi_dVd = T13 * dVth_dVd;
    dDenomi_dVb = T13 * dVth_dVb + p->uc * T4;
    dDenomi_dVg += T7;
  } else {
    T0 = (I->Vgsteff + I->vtfbphi1) / toxe;
    T1 = exp(p->eu * log(T0));
    dT1_dVg = T1 * p->eu / T0 / toxe;
    T2 = p->ua + p->uc * Vbseff;
    T3 = T0 / toxe;
    T12 = sqrt(Vth * Vth + 0.0001);
    T9 = 1.0 / (I->Vgsteff + 2 * T12);
    T10 = T9 * toxe;
    T8 = p->ud * T10 * T10 * Vth;
    T6 = T8 * Vth;
    T5 = T1 * T2 + T6;
    T7 = -2.0 * T6 * T9;
    T11 = T7 * Vth / T12;
    dDenomi_dVg = T2 * dT1_dVg + T7;
    T13 = 2.0 * (T11 + T8);
    dDenomi_dVd = T13 * dVth_dVd;
    dDenomi_dVb = T13 * dVth_dVb + T1 * p->uc;
  }
  if (T5 >= -0.8) {
    Denomi = 1.0 + T5;
  } else {
    T9 = 1.0 / (7.0 + 10.0 * T5);
    Denomi = (0.6 + T5) * T9;
    T9 *= T9;
    dDenomi_dVg *= T9;
    dDenomi_dVd *= T9;
    dDenomi_dVb *= T9;
  }
  I->ueff = I->u0temp / Denomi;
  T9 = -I->ueff / Denomi;
  dueff_dVg = T9 * dDenomi_dVg;
  dueff_dVd = T9 * dDenomi_dVd;
  dueff_dVb = T9 * dDenomi_dVb;

  /* saturation field and Vdsat (:9344-9450) */
  WVCox = Weff * I->vsattemp * m->coxe;
  WVCoxRds = WVCox * Rds;
  Esat = 2.0 * I->vsattemp / I->ueff;
  I->EsatL = Esat * Leff;
  T0 = -I->EsatL / I->ueff;
  dEsatL_dVg = T0 * dueff_dVg;
  dEsatL_dVd = T0 * dueff_dVd;
  dEsatL_dVb = T0 * dueff_dVb;
  a1 = p->a1;
  if (a1 == 0.0) {
    Lambda = p->a2;
    dLambda_dVg = 0.0;
  } else if (a1 > 0.0) {
    T0 = 1.0 - p->a2;
    T1 = T0 - p->a1 * I->Vgsteff - 0.0001;
    T2 = sqrt(T1 * T1 + 0.0004 * T0);
    Lambda = p->a2 + T0 - 0.5 * (T1 + T2);
    dLambda_dVg = 0.5 * p->a1 * (1.0 + T1 / T2);
  } else {
    T1 = p->a2 + p->a1 * I->Vgsteff - 0.0001;
    T2 = sqrt(T1 * T1 + 0.0004 * p->a2);
    Lambda = 0.5 * (T1 + T2);
    dLambda_dVg = 0.5 * p->a1 * (1.0 + T1 / T2);
  }
  Vgst2Vtm = I->Vgsteff + 2.0 * Vtm;
  if (Rds > 0) {
    tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff;
    tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff;
  } else {
    tmp2 = dWeff_dVg / Weff;
    tmp3 = dWeff_dVb / Weff;
  }
  if ((Rds == 0.0) && (Lambda == 1.0)) {
    T0 = 1.0 / (I->Abulk * I->EsatL + Vgst2Vtm);
    tmp1 = 0.0;
    T1 = T0 * T0;
    T2 = Vgst2Vtm * T0;
    T3 = I->EsatL * Vgst2Vtm;
    Vdsat = T3 * T0;
    dT0_dVg = -(I->Abulk * dEsatL_dVg + I->EsatL * dAbulk_dVg + 1.0) * T1;
    dT0_dVd = -(I->Abulk * dEsatL_dVd) * T1;
    dT0_dVb = -(I->Abulk * dEsatL_dVb + dAbulk_dVb * I->EsatL) * T1;
    dVdsat_dVg = T3 * dT0_dVg + T2 * dEsatL_dVg + I->EsatL * T0;
    dVdsat_dVd = T3 * dT0_dVd + T2 * dEsatL_dVd;
    dVdsat_dVb = T3 * dT0_dVb + T2 * dEsatL_dVb;
  } else {
    tmp1 = dLambda_dVg / (Lambda * Lambda);
    T9 = I->Abulk * WVCoxRds;
    T8 = I->Abulk * T9;
    T7 = Vgst2Vtm * T9;
    T6 = Vgst2Vtm * WVCoxRds;
    T0 = 2.0 * I->Abulk * (T9 - 1.0 + 1.0 / Lambda);
    dT0_dVg = 2.0 * (T8 * tmp2 - I->Abulk * tmp1 + (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
    dT0_dVb = 2.0 * (T8 * (2.0 / I->Abulk * dAbulk_dVb + tmp3) + (1.0 / Lambda - 1.0) * dAbulk_dVb);
    dT0_dVd = 0.0;
    T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + I->Abulk * I->EsatL + 3.0 * T7;
    dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + I->Abulk * dEsatL_dVg
              + I->EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
    dT1_dVb = I->Abulk * dEsatL_dVb + I->EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
    dT1_dVd = I->Abulk * dEsatL_dVd;
    T2 = Vgst2Vtm * (I->EsatL + 2.0 * T6);
    dT2_dVg = I->EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
    dT2_dVb = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
    dT2_dVd = Vgst2Vtm * dEsatL_dVd;
    T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
    Vdsat = (T1 - T3) / T0;
    dT3_dVg = (T1 * dT1_dVg - 2.0 * (T0 * dT2_dVg + T2 * dT0_dVg)) / T3;
    dT3_dVd = (T1 * dT1_dVd - 2.0 * (T0 * dT2_dVd + T2 * dT0_dVd)) / T3;
    dT3_dVb = (T1 * dT1_dVb - 2.0 * (T0 * dT2_dVb + T2 * dT0_dVb)) / T3;
    dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
    dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb) / T3 - Vdsat * dT0_dVb) / T0;
    dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
  }
  I->vdsat = Vdsat;

  /* effective Vds (:9453-9497) */
  T1 = Vdsat - Vds - p->delta;
  dT1_dVg = dVdsat_dVg;
  dT1_dVd = dVdsat_dVd - 1.0;
  dT1_dVb = dVdsat_dVb;
  T2 = sqrt(T1 * T1 + 4.0 * p->delta * Vdsat);
  T0 = T1 / T2;
  T9 = 2.0 * p->delta;
  T3 = T9 / T2;
  dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
  dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
  dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
  if (T1 >= 0.0) {
    Vdseff_ = Vdsat - 0.5 * (T1 + T2);
    dVdseff_dVg = dVdsat_dVg - 0.5 * (dT1_dVg + dT2_dVg);
    dVdseff_dVd = dVdsat_dVd - 0.5 * (dT1_dVd + dT2_dVd);
    dVdseff_dVb = dVdsat_dVb - 0.5 * (dT1_dVb + dT2_dVb);
  } else {
    T4 = T9 / (T2 - T1);
    T5 = 1.0 - T4;
    T6 = Vdsat * T4 / (T2 - T1);
    Vdseff_ = Vdsat * T5;
    dVdseff_dVg = dVdsat_dVg * T5 + T6 * (dT2_dVg - dT1_dVg);
    dVdseff_dVd = dVdsat_dVd * T5 + T6 * (dT2_dVd - dT1_dVd);
    dVdseff_dVb = dVdsat_dVb * T5 + T6 * (dT2_dVb - dT1_dVb);
  }
  if (Vds == 0.0) {
    Vdseff_ = 0.0;
    dVdseff_dVg = 0.0;
    dVdseff_dVb = 0.0;
  }
  if (Vdseff_ > Vds) Vdseff_ = Vds;
  diffVds = Vds - Vdseff_;
  I->Vdseff = Vdseff_;
  /* the DC values, kept for reading (the C-V section reuses these fields) */
  I->dcVgsteff = I->Vgsteff;
  I->dcAbulk = I->Abulk;
  I->dcVdseff = I->Vdseff;

  /* velocity overshoot (:9500-9540) */
  if ((MG(lambda)) && (M(lambda) > 0.0)) {
    T1 = Leff * I->ueff;
    T2 = p->lambda / T1;
    T3 = -T2 / T1 * Leff;
    dT2_dVd = T3 * dueff_dVd;
    dT2_dVg = T3 * dueff_dVg;
    dT2_dVb = T3 * dueff_dVb;
    T5 = 1.0 / (Esat * p->litl);
    T4 = -T5 / I->EsatL;
    dT5_dVg = dEsatL_dVg * T4;
    dT5_dVd = dEsatL_dVd * T4;
    dT5_dVb = dEsatL_dVb * T4;
    T6 = 1.0 + diffVds * T5;
    dT6_dVg = dT5_dVg * diffVds - dVdseff_dVg * T5;
    dT6_dVd = dT5_dVd * diffVds + (1.0 - dVdseff_dVd) * T5;
    dT6_dVb = dT5_dVb * diffVds - dVdseff_dVb * T5;
    T7 = 2.0 / (T6 * T6 + 1.0);
    T8 = 1.0 - T7;
    T9 = T6 * T7 * T7;
    dT8_dVg = T9 * dT6_dVg;
    dT8_dVd = T9 * dT6_dVd;
    dT8_dVb = T9 * dT6_dVb;
    T10 = 1.0 + T2 * T8;
    dT10_dVg = dT2_dVg * T8 + T2 * dT8_dVg;
    dT10_dVd = dT2_dVd * T8 + T2 * dT8_dVd;
    dT10_dVb = dT2_dVb * T8 + T2 * dT8_dVb;
    if (T10 == 1.0) dT10_dVg = dT10_dVd = dT10_dVb = 0.0;
    dEsatL_dVg *= T10;
    dEsatL_dVg += I->EsatL * dT10_dVg;
    dEsatL_dVd *= T10;
    dEsatL_dVd += I->EsatL * dT10_dVd;
    dEsatL_dVb *= T10;
    dEsatL_dVb += I->EsatL * dT10_dVb;
    I->EsatL *= T10;
  }

  /* Vasat (:9543-9565) */
  tmp4 = 1.0 - 0.5 * I->Abulk * Vdsat / Vgst2Vtm;
  T9 = WVCoxRds * I->Vgsteff;
  T8 = T9 / Vgst2Vtm;
  T0 = I->EsatL + Vdsat + 2.0 * T9 * tmp4;
  T7 = 2.0 * WVCoxRds * tmp4;
  dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * I->Vgsteff)
            - T8 * (I->Abulk * dVdsat_dVg - I->Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
  dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * I->Vgsteff
            - T8 * (dAbulk_dVb * Vdsat + I->Abulk * dVdsat_dVb);
  dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * I->Abulk * dVdsat_dVd;
  T9 = WVCoxRds * I->Abulk;
  T1 = 2.0 / Lambda - 1.0 + T9;
  dT1_dVg = -2.0 * tmp1 + WVCoxRds * (I->Abulk * tmp2 + dAbulk_dVg);
  dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
  Vasat = T0 / T1;
  dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
  dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
  dVasat_dVd = dT0_dVd / T1;

  /* effective oxide capacitance, beta, Idl (:9568-9622) */
  tmp1 = I->vtfbphi2;
  tmp2 = 2.0e8 * M(toxp);
  dT0_dVg = 1.0 / tmp2;
  T0 = (I->Vgsteff + tmp1) * dT0_dVg;
  tmp3 = exp(M(bdos) * 0.7 * log(T0));
  T1 = 1.0 + tmp3;
  T2 = M(bdos) * 0.7 * tmp3 / T0;
  Tcen = M(ados) * 1.9e-9 / T1;
  dTcen_dVg = -Tcen * T2 * dT0_dVg / T1;
  Coxeff = epssub * m->coxp / (epssub + m->coxp * Tcen);
  dCoxeff_dVg = -Coxeff * Coxeff * dTcen_dVg / epssub;
  CoxeffWovL = Coxeff * Weff / Leff;
  beta = I->ueff * CoxeffWovL;
  T3 = I->ueff / Leff;
  dbeta_dVg = CoxeffWovL * dueff_dVg + T3 * (Weff * dCoxeff_dVg + Coxeff * dWeff_dVg);
  dbeta_dVd = CoxeffWovL * dueff_dVd;
  dbeta_dVb = CoxeffWovL * dueff_dVb + T3 * Coxeff * dWeff_dVb;
  I->AbovVgst2Vtm = I->Abulk / Vgst2Vtm;
  T0 = 1.0 - 0.5 * Vdseff_ * I->AbovVgst2Vtm;
  dT0_dVg = -0.5 * (I->Abulk * dVdseff_dVg - I->Abulk * Vdseff_ / Vgst2Vtm + Vdseff_ * dAbulk_dVg) / Vgst2Vtm;
  dT0_dVd = -0.5 * I->Abulk * dVdseff_dVd / Vgst2Vtm;
  dT0_dVb = -0.5 * (I->Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff_) / Vgst2Vtm;
  fgche1 = I->Vgsteff * T0;
  dfgche1_dVg = I->Vgsteff * dT0_dVg + T0;
  dfgche1_dVd = I->Vgsteff * dT0_dVd;
  dfgche1_dVb = I->Vgsteff * dT0_dVb;
  T9 = Vdseff_ / I->EsatL;
  fgche2 = 1.0 + T9;
  dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / I->EsatL;
  dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / I->EsatL;
  dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / I->EsatL;
  gche = beta * fgche1 / fgche2;
  dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
  dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
  dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
  T0 = 1.0 + gche * Rds;
  Idl = gche / T0;
  T1 = (1.0 - Idl * Rds) / T0;
  T2 = Idl * Idl;
  dIdl_dVg = T1 * dgche_dVg - T2 * dRds_dVg;
  dIdl_dVd = T1 * dgche_dVd;
  dIdl_dVb = T1 * dgche_dVb - T2 * dRds_dVb;

  /* output resistance: CLM, DIBL, DITS, SCBE (:9625-9832) */
  if (p->fprout <= 0.0) {
    FP = 1.0;
    dFP_dVg = 0.0;
  } else {
    T9 = p->fprout * sqrt(Leff) / Vgst2Vtm;
    FP = 1.0 / (1.0 + T9);
    dFP_dVg = FP * FP * T9 / Vgst2Vtm;
  }
  T8 = p->pvag / I->EsatL;
  T9 = T8 * I->Vgsteff;
  if (T9 > -0.9) {
    PvagTerm = 1.0 + T9;
    dPvagTerm_dVg = T8 * (1.0 - I->Vgsteff * dEsatL_dVg / I->EsatL);
    dPvagTerm_dVb = -T9 * dEsatL_dVb / I->EsatL;
    dPvagTerm_dVd = -T9 * dEsatL_dVd / I->EsatL;
  } else {
    T4 = 1.0 / (17.0 + 20.0 * T9);
    PvagTerm = (0.8 + T9) * T4;
    T4 *= T4;
    dPvagTerm_dVg = T8 * (1.0 - I->Vgsteff * dEsatL_dVg / I->EsatL) * T4;
    T9 *= T4 / I->EsatL;
    dPvagTerm_dVb = -T9 * dEsatL_dVb;
    dPvagTerm_dVd = -T9 * dEsatL_dVd;
  }
  if ((p->pclm > OR_MIN_EXP) && (diffVds > 1.0e-10)) {
    T0 = 1.0 + Rds * Idl;
    dT0_dVg = dRds_dVg * Idl + Rds * dIdl_dVg;
    dT0_dVd = Rds * dIdl_dVd;
    dT0_dVb = dRds_dVb * Idl + Rds * dIdl_dVb;
    T2 = Vdsat / Esat;
    T1 = Leff + T2;
    dT1_dVg = (dVdsat_dVg - T2 * dEsatL_dVg / Leff) / Esat;
    dT1_dVd = (dVdsat_dVd - T2 * dEsatL_dVd / Leff) / Esat;
    dT1_dVb = (dVdsat_dVb - T2 * dEsatL_dVb / Leff) / Esat;
    Cclm = FP * PvagTerm * T0 * T1 / (p->pclm * p->litl);
    dCclm_dVg = Cclm * (dFP_dVg / FP + dPvagTerm_dVg / PvagTerm + dT0_dVg / T0 + dT1_dVg / T1);
    dCclm_dVb = Cclm * (dPvagTerm_dVb / PvagTerm + dT0_dVb / T0 + dT1_dVb / T1);
    dCclm_dVd = Cclm * (dPvagTerm_dVd / PvagTerm + dT0_dVd / T0 + dT1_dVd / T1);
    VACLM = Cclm * diffVds;
    dVACLM_dVg = dCclm_dVg * diffVds - dVdseff_dVg * Cclm;
    dVACLM_dVb = dCclm_dVb * diffVds - dVdseff_dVb * Cclm;
    dVACLM_dVd = dCclm_dVd * diffVds + (1.0 - dVdseff_dVd) * Cclm;
  } else {
    VACLM = Cclm = OR_MAX_EXP;
    dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
    dCclm_dVd = dCclm_dVg = dCclm_dVb = 0.0;
  }
  if (p->thetaRout > OR_MIN_EXP) {
    T8 = I->Abulk * Vdsat;
    T0 = Vgst2Vtm * T8;
    dT0_dVg = Vgst2Vtm * I->Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
    dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + I->Abulk * dVdsat_dVb);
    dT0_dVd = Vgst2Vtm * I->Abulk * dVdsat_dVd;
    T1 = Vgst2Vtm + T8;
    dT1_dVg = 1.0 + I->Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
    dT1_dVb = I->Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
    dT1_dVd = I->Abulk * dVdsat_dVd;
    T9 = T1 * T1;
    T2 = p->thetaRout;
    VADIBL = (Vgst2Vtm - T0 / T1) / T2;
    dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
    dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
    dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
    T7 = p->pdiblb * Vbseff;
    if (T7 >= -0.9) {
      T3 = 1.0 / (1.0 + T7);
      VADIBL *= T3;
      dVADIBL_dVg *= T3;
      dVADIBL_dVb = (dVADIBL_dVb - VADIBL * p->pdiblb) * T3;
      dVADIBL_dVd *= T3;
    } else {
      T4 = 1.0 / (0.8 + T7);
      T3 = (17.0 + 20.0 * T7) * T4;
      dVADIBL_dVg *= T3;
      dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * p->pdiblb * T4 * T4;
      dVADIBL_dVd *= T3;
      VADIBL *= T3;
    }
    dVADIBL_dVg = dVADIBL_dVg * PvagTerm + VADIBL * dPvagTerm_dVg;
    dVADIBL_dVb = dVADIBL_dVb * PvagTerm + VADIBL * dPvagTerm_dVb;
    dVADIBL_dVd = dVADIBL_dVd * PvagTerm + VADIBL * dPvagTerm_dVd;
    VADIBL *= PvagTerm;
  } else {
    VADIBL = OR_MAX_EXP;
    dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
  }
  Va = Vasat + VACLM;
  dVa_dVg = dVasat_dVg + dVACLM_dVg;
  dVa_dVb = dVasat_dVb + dVACLM_dVb;
  dVa_dVd = dVasat_dVd + dVACLM_dVd;
  T0 = p->pditsd * Vds;
  if (T0 > OR_EXP_THRESHOLD) {
    T1 = OR_MAX_EXP;
    dT1_dVd = 0;
  } else {
    T1 = exp(T0);
    dT1_dVd = T1 * p->pditsd;
  }
  if (p->pdits > OR_MIN_EXP) {
    T2 = 1.0 + M(pditsl) * Leff;
    VADITS = (1.0 + T2 * T1) / p->pdits;
    dVADITS_dVg = VADITS * dFP_dVg;
    dVADITS_dVd = FP * T2 * dT1_dVd / p->pdits;
    VADITS *= FP;
  } else {
    VADITS = OR_MAX_EXP;
    dVADITS_dVg = dVADITS_dVd = 0;
  }
  if (p->pscbe2 > 0.0) {
    if (diffVds > p->pscbe1 * p->litl / OR_EXP_THRESHOLD) {
      T0 = p->pscbe1 * p->litl / diffVds;
      VASCBE = Leff * exp(T0) / p->pscbe2;
      T1 = T0 * VASCBE / diffVds;
      dVASCBE_dVg = T1 * dVdseff_dVg;
      dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
      dVASCBE_dVb = T1 * dVdseff_dVb;
    } else {
      VASCBE = OR_MAX_EXP * Leff / p->pscbe2;
      dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
    }
  } else {
    VASCBE = OR_MAX_EXP;
    dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
  }
  T9 = diffVds / VADIBL;
  T0 = 1.0 + T9;
  Idsa = Idl * T0;
  dIdsa_dVg = T0 * dIdl_dVg - Idl * (dVdseff_dVg + T9 * dVADIBL_dVg) / VADIBL;
  dIdsa_dVd = T0 * dIdl_dVd + Idl * (1.0 - dVdseff_dVd - T9 * dVADIBL_dVd) / VADIBL;
  dIdsa_dVb = T0 * dIdl_dVb - Idl * (dVdseff_dVb + T9 * dVADIBL_dVb) / VADIBL;
  T9 = diffVds / VADITS;
  T0 = 1.0 + T9;
  dIdsa_dVg = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVADITS_dVg) / VADITS;
  dIdsa_dVd = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVADITS_dVd) / VADITS;
  dIdsa_dVb = T0 * dIdsa_dVb - Idsa * dVdseff_dVb / VADITS;
  Idsa *= T0;
  T0 = log(Va / Vasat);
  dT0_dVg = dVa_dVg / Va - dVasat_dVg / Vasat;
  dT0_dVb = dVa_dVb / Va - dVasat_dVb / Vasat;
  dT0_dVd = dVa_dVd / Va - dVasat_dVd / Vasat;
  T1 = T0 / Cclm;
  T9 = 1.0 + T1;
  dT9_dVg = (dT0_dVg - T1 * dCclm_dVg) / Cclm;
  dT9_dVb = (dT0_dVb - T1 * dCclm_dVb) / Cclm;
  dT9_dVd = (dT0_dVd - T1 * dCclm_dVd) / Cclm;
  dIdsa_dVg = dIdsa_dVg * T9 + Idsa * dT9_dVg;
  dIdsa_dVb = dIdsa_dVb * T9 + Idsa * dT9_dVb;
  dIdsa_dVd = dIdsa_dVd * T9 + Idsa * dT9_dVd;
  Idsa *= T9;

  /* substrate current (:9835-9877) */
  tmp = p->alpha0 + p->alpha1 * Leff;
  if ((tmp <= 0.0) || (p->beta0 <= 0.0)) {
    Isub = Gbd = Gbb = Gbg = 0.0;
  } else {
    T2 = tmp / Leff;
    if (diffVds > p->beta0 / OR_EXP_THRESHOLD) {
      T0 = -p->beta0 / diffVds;
      T1 = T2 * diffVds * exp(T0);
      T3 = T1 / diffVds * (T0 - 1.0);
      dT1_dVg = T3 * dVdseff_dVg;
      dT1_dVd = T3 * (dVdseff_dVd - 1.0);
      dT1_dVb = T3 * dVdseff_dVb;
    } else {
      T3 = T2 * OR_MIN_EXP;
      T1 = T3 * diffVds;
      dT1_dVg = -T3 * dVdseff_dVg;
      dT1_dVd = T3 * (1.0 - dVdseff_dVd);
      dT1_dVb = -T3 * dVdseff_dVb;
    }
    T4 = Idsa * Vdseff_;
    Isub = T1 * T4;
    Gbg = T1 * (dIdsa_dVg * Vdseff_ + Idsa * dVdseff_dVg) + T4 * dT1_dVg;
    Gbd = T1 * (dIdsa_dVd * Vdseff_ + Idsa * dVdseff_dVd) + T4 * dT1_dVd;
    Gbb = T1 * (dIdsa_dVb * Vdseff_ + Idsa * dVdseff_dVb) + T4 * dT1_dVb;
    Gbd += Gbg * dVgsteff_dVd;
    Gbb += Gbg * dVgsteff_dVb;
    Gbg *= dVgsteff_dVg;
    Gbb *= dVbseff_dVb;
  }
  I->csub = Isub;
  I->gbbs = Gbb;
  I->gbgs = Gbg;
  I->gbds = Gbd;

  /* drain current with SCBE (:9880-9902) */
  T9 = diffVds / VASCBE;
  T0 = 1.0 + T9;
  Ids = Idsa * T0;
  Gm_ = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
  Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
  Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
  tmp1 = Gds + Gm_ * dVgsteff_dVd;
  tmp2 = Gmb + Gm_ * dVgsteff_dVb;
  tmp3 = Gm_;
  Gm_ = (Ids * dVdseff_dVg + Vdseff_ * tmp3) * dVgsteff_dVg;
  Gds = Ids * (dVdseff_dVd + dVdseff_dVg * dVgsteff_dVd) + Vdseff_ * tmp1;
  Gmb = (Ids * (dVdseff_dVb + dVdseff_dVg * dVgsteff_dVb) + Vdseff_ * tmp2) * dVbseff_dVb;
  I->cdrain = Ids * Vdseff_;

  /* source-end velocity limit (:9905-9945) */
  if ((MG(vtl)) && (M(vtl) > 0.0)) {
    T12 = 1.0 / Leff / CoxeffWovL;
    T11 = T12 / I->Vgsteff;
    T10 = -T11 / I->Vgsteff;
    vs = I->cdrain * T11;
    dvs_dVg = Gm_ * T11 + I->cdrain * T10 * dVgsteff_dVg;
    dvs_dVd = Gds * T11 + I->cdrain * T10 * dVgsteff_dVd;
    dvs_dVb = Gmb * T11 + I->cdrain * T10 * dVgsteff_dVb;
    T0 = 2 * B4_MM;
    T1 = vs / (p->vtl * p->tfactor);
    if (T1 > 0.0) {
      T2 = 1.0 + exp(T0 * log(T1));
      T3 = (T2 - 1.0) * T0 / vs;
      Fsevl = 1.0 / exp(log(T2) / T0);
      dT2_dVg = T3 * dvs_dVg;
      dT2_dVd = T3 * dvs_dVd;
      dT2_dVb = T3 * dvs_dVb;
      T4 = -1.0 / T0 * Fsevl / T2;
      dFsevl_dVg = T4 * dT2_dVg;
      dFsevl_dVd = T4 * dT2_dVd;
      dFsevl_dVb = T4 * dT2_dVb;
    } else {
      Fsevl = 1.0;
      dFsevl_dVg = 0.0;
      dFsevl_dVd = 0.0;
      dFsevl_dVb = 0.0;
    }
    Gm_ *= Fsevl;
    Gm_ += I->cdrain * dFsevl_dVg;
    Gmb *= Fsevl;
    Gmb += I->cdrain * dFsevl_dVb;
    Gds *= Fsevl;
    Gds += I->cdrain * dFsevl_dVd;
    I->cdrain *= Fsevl;
  }
  I->gds = Gds;
  I->gm = Gm_;
  I->gmbs = Gmb;
  I->IdovVds = Ids;
  if (I->IdovVds <= 1.0e-9) I->IdovVds = 1.0e-9;
  /* gate resistance conductances (:9953-9988); without rgateMod > 1 or NQS
     they stay at the zeros setupFVectorVars left */
  if ((I->irgateMod > 1) || (I->itrnqsMod != 0) || (I->iacnqsMod != 0)) {
    T9 = p->xrcrg2 * m->vtm;
    T0 = T9 * beta;
    dT0_dVd = (dbeta_dVd + dbeta_dVg * dVgsteff_dVd) * T9;
    dT0_dVb = (dbeta_dVb + dbeta_dVg * dVgsteff_dVb) * T9;
    dT0_dVg = dbeta_dVg * T9;
    I->gcrg = p->xrcrg1 * (T0 + Ids);
    I->gcrgd = p->xrcrg1 * (dT0_dVd + tmp1);
    I->gcrgb = p->xrcrg1 * (dT0_dVb + tmp2) * dVbseff_dVb;
    I->gcrgg = p->xrcrg1 * (dT0_dVg + tmp3) * dVgsteff_dVg;
    if (nf != 1.0) {
      I->gcrg *= nf;
      I->gcrgg *= nf;
      I->gcrgd *= nf;
      I->gcrgb *= nf;
    }
    if (I->irgateMod == 2) {
      T10 = I->grgeltd * I->grgeltd;
      T11 = I->grgeltd + I->gcrg;
      I->gcrg = I->grgeltd * I->gcrg / T11;
      T12 = T10 / T11 / T11;
      I->gcrgg *= T12;
      I->gcrgd *= T12;
      I->gcrgb *= T12;
    }
    I->gcrgs = -(I->gcrgg + I->gcrgd + I->gcrgb);
  } else {
    I->gcrg = I->gcrgd = I->gcrgg = I->gcrgs = I->gcrgb = 0.0;
  }
  /* bias-dependent external source/drain resistance (:9991-10069) */
  if (M(rdsMod)) {
    T0 = I->vgs - p->vfbsd;
    T1 = sqrt(T0 * T0 + 1.0e-4);
    I->vgs_eff = 0.5 * (T0 + T1);
    I->dvgs_eff_dvg = I->vgs_eff / T1;
    T0 = 1.0 + p->prwg * I->vgs_eff;
    dT0_dvg = -p->prwg / T0 / T0 * I->dvgs_eff_dvg;
    T1 = -p->prwb * I->vbs;
    dT1_dvb = -p->prwb;
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    dT3_dvg = T3 / (T3 - T2);
    dT3_dvb = dT3_dvg * dT1_dvb;
    dT3_dvg *= dT0_dvg;
    T4 = p->rs0 * 0.5;
    Rs = p->rswmin + T3 * T4;
    dRs_dvg = T4 * dT3_dvg;
    dRs_dvb = T4 * dT3_dvb;
    T0 = 1.0 + I->sourceConductance * Rs;
    I->gstot = I->sourceConductance / T0;
    T0 = -I->gstot * I->gstot;
    dgstot_dvd = 0.0;
    dgstot_dvg = T0 * dRs_dvg;
    dgstot_dvb = T0 * dRs_dvb;
    dgstot_dvs = -(dgstot_dvg + dgstot_dvb + dgstot_dvd);
    T0 = I->vgd - p->vfbsd;
    T1 = sqrt(T0 * T0 + 1.0e-4);
    I->vgd_eff = 0.5 * (T0 + T1);
    I->dvgd_eff_dvg = I->vgd_eff / T1;
    T0 = 1.0 + p->prwg * I->vgd_eff;
    dT0_dvg = -p->prwg / T0 / T0 * I->dvgd_eff_dvg;
    T1 = -p->prwb * I->vbd;
    dT1_dvb = -p->prwb;
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    dT3_dvg = T3 / (T3 - T2);
    dT3_dvb = dT3_dvg * dT1_dvb;
    dT3_dvg *= dT0_dvg;
    T4 = p->rd0 * 0.5;
    Rd = p->rdwmin + T3 * T4;
    dRd_dvg = T4 * dT3_dvg;
    dRd_dvb = T4 * dT3_dvb;
    T0 = 1.0 + I->drainConductance * Rd;
    I->gdtot = I->drainConductance / T0;
    T0 = -I->gdtot * I->gdtot;
    dgdtot_dvs = 0.0;
    dgdtot_dvg = T0 * dRd_dvg;
    dgdtot_dvb = T0 * dRd_dvb;
    dgdtot_dvd = -(dgdtot_dvg + dgdtot_dvb + dgdtot_dvs);
    I->gstotd = I->vses * dgstot_dvd;
    I->gstotg = I->vses * dgstot_dvg;
    I->gstots = I->vses * dgstot_dvs;
    I->gstotb = I->vses * dgstot_dvb;
    T2 = I->vdes - I->vds;
    I->gdtotd = T2 * dgdtot_dvd;
    I->gdtotg = T2 * dgdtot_dvg;
    I->gdtots = T2 * dgdtot_dvs;
    I->gdtotb = T2 * dgdtot_dvb;
  } else {
    I->gstot = I->gstotd = I->gstotg = 0.0;
    I->gstots = I->gstotb = 0.0;
    I->gdtot = I->gdtotd = I->gdtotg = 0.0;
    I->gdtots = I->gdtotb = 0.0;
  }

  /* GIDL (:10074-10126) */
  if (M(mtrlMod) == 0) T0 = 3.0 * toxe;
  else T0 = M(epsrsub) * toxe / epsrox;
  if (M(mtrlMod) == 0) T1 = (I->vds - I->vgs_eff - p->egidl) / T0;
  else T1 = (I->vds - I->vgs_eff - p->egidl + p->vfbsd) / T0;
  if ((p->agidl <= 0.0) || (p->bgidl <= 0.0) || (T1 <= 0.0) || (p->cgidl <= 0.0) || (I->vbd > 0.0)) {
    I->Igidl = Ggidld = Ggidlg = Ggidlb = 0.0;
  } else {
    dT1_dVd = 1.0 / T0;
    dT1_dVg = -I->dvgs_eff_dvg * dT1_dVd;
    T2 = p->bgidl / T1;
    if (T2 < 100.0) {
      I->Igidl = p->agidl * p->weffCJ * T1 * exp(-T2);
      T3 = I->Igidl * (1.0 + T2) / T1;
      Ggidld = T3 * dT1_dVd;
      Ggidlg = T3 * dT1_dVg;
    } else {
      I->Igidl = p->agidl * p->weffCJ * 3.720075976e-44;
      Ggidld = I->Igidl * dT1_dVd;
      Ggidlg = I->Igidl * dT1_dVg;
      I->Igidl *= T1;
    }
    T4 = I->vbd * I->vbd;
    T5 = -I->vbd * T4;
    T6 = p->cgidl + T5;
    T7 = T5 / T6;
    T8 = 3.0 * p->cgidl * T4 / T6 / T6;
    Ggidld = Ggidld * T7 + I->Igidl * T8;
    Ggidlg = Ggidlg * T7;
    Ggidlb = -I->Igidl * T8;
    I->Igidl *= T7;
  }
  I->ggidld = Ggidld;
  I->ggidlg = Ggidlg;
  I->ggidlb = Ggidlb;
  /* GISL (:10132-10179) */
  if (M(mtrlMod) == 0) T1 = (-I->vds - I->vgd_eff - p->egisl) / T0;
  else T1 = (-I->vds - I->vgd_eff - p->egisl + p->vfbsd) / T0;
  if ((p->agisl <= 0.0) || (p->bgisl <= 0.0) || (T1 <= 0.0) || (p->cgisl <= 0.0) || (I->vbs > 0.0)) {
    I->Igisl = Ggisls = Ggislg = Ggislb = 0.0;
  } else {
    dT1_dVd = 1.0 / T0;
    dT1_dVg = -I->dvgd_eff_dvg * dT1_dVd;
    T2 = p->bgisl / T1;
    if (T2 < 100.0) {
      I->Igisl = p->agisl * p->weffCJ * T1 * exp(-T2);
      T3 = I->Igisl * (1.0 + T2) / T1;
      Ggisls = T3 * dT1_dVd;
      Ggislg = T3 * dT1_dVg;
    } else {
      I->Igisl = p->agisl * p->weffCJ * 3.720075976e-44;
      Ggisls = I->Igisl * dT1_dVd;
      Ggislg = I->Igisl * dT1_dVg;
      I->Igisl *= T1;
    }
    T4 = I->vbs * I->vbs;
    T5 = -I->vbs * T4;
    T6 = p->cgisl + T5;
    T7 = T5 / T6;
    T8 = 3.0 * p->cgisl * T4 / T6 / T6;
    Ggisls = Ggisls * T7 + I->Igisl * T8;
    Ggislg = Ggislg * T7;
    Ggislb = -I->Igisl * T8;
    I->Igisl *= T7;
  }
  I->ggisls = Ggisls;
  I->ggislg = Ggislg;
  I->ggislb = Ggislb;
  /* gate tunnelling current (:10182-10621) */
  Vfb = 0.0;
  Voxacc = dVoxacc_dVg = dVoxacc_dVb = 0.0;
  Voxdepinv = dVoxdepinv_dVg = dVoxdepinv_dVd = dVoxdepinv_dVb = 0.0;
  if ((M(igcMod) != 0) || (M(igbMod) != 0)) {
    Vfb = I->vfbzb;
    V3 = Vfb - Vgs_eff + Vbseff - B4_DELTA_3;
    if (Vfb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * B4_DELTA_3 * Vfb);
    else T0 = sqrt(V3 * V3 + 4.0 * B4_DELTA_3 * Vfb);
    T1 = 0.5 * (1.0 + V3 / T0);
    Vfbeff = Vfb - 0.5 * (V3 + T0);
    dVfbeff_dVg = T1 * dVgs_eff_dVg;
    dVfbeff_dVb = -T1;
    Voxacc = Vfb - Vfbeff;
    dVoxacc_dVg = -dVfbeff_dVg;
    dVoxacc_dVb = -dVfbeff_dVb;
    if (Voxacc < 0.0) Voxacc = dVoxacc_dVg = dVoxacc_dVb = 0.0;
    T0 = 0.5 * p->k1ox;
    T3 = Vgs_eff - Vfbeff - Vbseff - I->Vgsteff;
    if (p->k1ox == 0.0) {
      Voxdepinv = dVoxdepinv_dVg = dVoxdepinv_dVd = dVoxdepinv_dVb = 0.0;
    } else if (T3 < 0.0) {
      Voxdepinv = -T3;
      dVoxdepinv_dVg = -dVgs_eff_dVg + dVfbeff_dVg + dVgsteff_dVg;
      dVoxdepinv_dVd = dVgsteff_dVd;
      dVoxdepinv_dVb = dVfbeff_dVb + 1.0 + dVgsteff_dVb;
    } else {
      T1 = sqrt(T0 * T0 + T3);
      T2 = T0 / T1;
      Voxdepinv = p->k1ox * (T1 - T0);
      dVoxdepinv_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg);
      dVoxdepinv_dVd = -T2 * dVgsteff_dVd;
      dVoxdepinv_dVb = -T2 * (dVfbeff_dVb + 1.0 + dVgsteff_dVb);
    }
    Voxdepinv += I->Vgsteff;
    dVoxdepinv_dVg += dVgsteff_dVg;
    dVoxdepinv_dVd += dVgsteff_dVd;
    dVoxdepinv_dVb += dVgsteff_dVb;
  }
  if (M(tempMod) < 2) tmp = Vtm;
  else tmp = Vtm0;
  if (M(igcMod)) {
    T0 = tmp * p->nigc;
    VxNVt = 0.0;
    Vaux = dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
    if (M(igcMod) == 1) {
      VxNVt = (Vgs_eff - m->dtype * I->vth0) / T0;
      if (VxNVt > OR_EXP_THRESHOLD) {
        Vaux = Vgs_eff - m->dtype * I->vth0;
        dVaux_dVg = dVgs_eff_dVg;
        dVaux_dVd = 0.0;
        dVaux_dVb = 0.0;
      }
    } else if (M(igcMod) == 2) {
      VxNVt = (Vgs_eff - I->von) / T0;
      if (VxNVt > OR_EXP_THRESHOLD) {
        Vaux = Vgs_eff - I->von;
        dVaux_dVg = dVgs_eff_dVg;
        dVaux_dVd = -dVth_dVd;
        dVaux_dVb = -dVth_dVb;
      }
    }
    if (VxNVt < -OR_EXP_THRESHOLD) {
      Vaux = T0 * log(1.0 + OR_MIN_EXP);
      dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
    } else if ((VxNVt >= -OR_EXP_THRESHOLD) && (VxNVt <= OR_EXP_THRESHOLD)) {
      ExpVxNVt = exp(VxNVt);
      Vaux = T0 * log(1.0 + ExpVxNVt);
      dVaux_dVg = ExpVxNVt / (1.0 + ExpVxNVt);
      if (M(igcMod) == 1) {
        dVaux_dVd = 0.0;
        dVaux_dVb = 0.0;
      } else if (M(igcMod) == 2) {
        dVaux_dVd = -dVgs_eff_dVg * dVth_dVd;
        dVaux_dVb = -dVgs_eff_dVg * dVth_dVb;
      }
      dVaux_dVg *= dVgs_eff_dVg;
    }
    T2 = Vgs_eff * Vaux;
    dT2_dVg = dVgs_eff_dVg * Vaux + Vgs_eff * dVaux_dVg;
    dT2_dVd = Vgs_eff * dVaux_dVd;
    dT2_dVb = Vgs_eff * dVaux_dVb;
    T11 = p->Aechvb;
    T12 = p->Bechvb;
    T3 = p->aigc * p->cigc - p->bigc;
    T4 = p->bigc * p->cigc;
    T5 = T12 * (p->aigc + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
    if (T5 > OR_EXP_THRESHOLD) {
      T6 = OR_MAX_EXP;
      dT6_dVg = dT6_dVd = dT6_dVb = 0.0;
    } else if (T5 < -OR_EXP_THRESHOLD) {
      T6 = OR_MIN_EXP;
      dT6_dVg = dT6_dVd = dT6_dVb = 0.0;
    } else {
      T6 = exp(T5);
      dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * Voxdepinv);
      dT6_dVd = dT6_dVg * dVoxdepinv_dVd;
      dT6_dVb = dT6_dVg * dVoxdepinv_dVb;
      dT6_dVg *= dVoxdepinv_dVg;
    }
    Igc = T11 * T2 * T6;
    dIgc_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
    dIgc_dVd = T11 * (T2 * dT6_dVd + T6 * dT2_dVd);
    dIgc_dVb = T11 * (T2 * dT6_dVb + T6 * dT2_dVb);
    if (MG(pigcd)) {
      Pigcd = p->pigcd;
      dPigcd_dVg = dPigcd_dVd = dPigcd_dVb = 0.0;
    } else {
      T11 = p->Bechvb * toxe;
      T12 = I->Vgsteff + 1.0e-20;
      T13 = T11 / T12 / T12;
      T14 = -T13 / T12;
      Pigcd = T13 * (1.0 - 0.5 * Vdseff_ / T12);
      dPigcd_dVg = T14 * (2.0 + 0.5 * (dVdseff_dVg - 3.0 * Vdseff_ / T12));
      dPigcd_dVd = 0.5 * T14 * dVdseff_dVd;
      dPigcd_dVb = 0.5 * T14 * dVdseff_dVb;
    }
    T7 = -Pigcd * Vdseff_;
    dT7_dVg = -Vdseff_ * dPigcd_dVg - Pigcd * dVdseff_dVg;
    dT7_dVd = -Vdseff_ * dPigcd_dVd - Pigcd * dVdseff_dVd + dT7_dVg * dVgsteff_dVd;
    dT7_dVb = -Vdseff_ * dPigcd_dVb - Pigcd * dVdseff_dVb + dT7_dVg * dVgsteff_dVb;
    dT7_dVg *= dVgsteff_dVg;
    dT7_dVb *= dVbseff_dVb;
    T8 = T7 * T7 + 2.0e-4;
    dT8_dVg = 2.0 * T7;
    dT8_dVd = dT8_dVg * dT7_dVd;
    dT8_dVb = dT8_dVg * dT7_dVb;
    dT8_dVg *= dT7_dVg;
    if (T7 > OR_EXP_THRESHOLD) {
      T9 = OR_MAX_EXP;
      dT9_dVg = dT9_dVd = dT9_dVb = 0.0;
    } else if (T7 < -OR_EXP_THRESHOLD) {
      T9 = OR_MIN_EXP;
      dT9_dVg = dT9_dVd = dT9_dVb = 0.0;
    } else {
      T9 = exp(T7);
      dT9_dVg = T9 * dT7_dVg;
      dT9_dVd = T9 * dT7_dVd;
      dT9_dVb = T9 * dT7_dVb;
    }
    T0 = T8 * T8;
    T1 = T9 - 1.0 + 1.0e-4;
    T10 = (T1 - T7) / T8;
    dT10_dVg = (dT9_dVg - dT7_dVg - T10 * dT8_dVg) / T8;
    dT10_dVd = (dT9_dVd - dT7_dVd - T10 * dT8_dVd) / T8;
    dT10_dVb = (dT9_dVb - dT7_dVb - T10 * dT8_dVb) / T8;
    I->Igcs = Igc * T10;
    dIgcs_dVg = dIgc_dVg * T10 + Igc * dT10_dVg;
    dIgcs_dVd = dIgc_dVd * T10 + Igc * dT10_dVd;
    dIgcs_dVb = dIgc_dVb * T10 + Igc * dT10_dVb;
    T1 = T9 - 1.0 - 1.0e-4;
    T10 = (T7 * T9 - T1) / T8;
    dT10_dVg = (dT7_dVg * T9 + (T7 - 1.0) * dT9_dVg - T10 * dT8_dVg) / T8;
    dT10_dVd = (dT7_dVd * T9 + (T7 - 1.0) * dT9_dVd - T10 * dT8_dVd) / T8;
    dT10_dVb = (dT7_dVb * T9 + (T7 - 1.0) * dT9_dVb - T10 * dT8_dVb) / T8;
    I->Igcd = Igc * T10;
    dIgcd_dVg = dIgc_dVg * T10 + Igc * dT10_dVg;
    dIgcd_dVd = dIgc_dVd * T10 + Igc * dT10_dVd;
    dIgcd_dVb = dIgc_dVb * T10 + Igc * dT10_dVb;
    I->gIgcsg = dIgcs_dVg;
    I->gIgcsd = dIgcs_dVd;
    I->gIgcsb = dIgcs_dVb * dVbseff_dVb;
    I->gIgcdg = dIgcd_dVg;
    I->gIgcdd = dIgcd_dVd;
    I->gIgcdb = dIgcd_dVb * dVbseff_dVb;
    /* source/drain edge tunnelling: rewrites the vgs_eff / vgd_eff members */
    T0 = I->vgs - (p->vfbsd + p->vfbsdoff);
    I->vgs_eff = sqrt(T0 * T0 + 1.0e-4);
    I->dvgs_eff_dvg = T0 / I->vgs_eff;
    T2 = I->vgs * I->vgs_eff;
    dT2_dVg = I->vgs * I->dvgs_eff_dvg + I->vgs_eff;
    T11 = p->AechvbEdgeS;
    T12 = p->BechvbEdge;
    T3 = p->aigs * p->cigs - p->bigs;
    T4 = p->bigs * p->cigs;
    T5 = T12 * (p->aigs + T3 * I->vgs_eff - T4 * I->vgs_eff * I->vgs_eff);
    if (T5 > OR_EXP_THRESHOLD) {
      T6 = OR_MAX_EXP;
      dT6_dVg = 0.0;
    } else if (T5 < -OR_EXP_THRESHOLD) {
      T6 = OR_MIN_EXP;
      dT6_dVg = 0.0;
    } else {
      T6 = exp(T5);
      dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * I->vgs_eff) * I->dvgs_eff_dvg;
    }
    I->Igs = T11 * T2 * T6;
    dIgs_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
    dIgs_dVs = -dIgs_dVg;
    T0 = I->vgd - (p->vfbsd + p->vfbsdoff);
    I->vgd_eff = sqrt(T0 * T0 + 1.0e-4);
    I->dvgd_eff_dvg = T0 / I->vgd_eff;
    T2 = I->vgd * I->vgd_eff;
    dT2_dVg = I->vgd * I->dvgd_eff_dvg + I->vgd_eff;
    T11 = p->AechvbEdgeD;
    T3 = p->aigd * p->cigd - p->bigd;
    T4 = p->bigd * p->cigd;
    T5 = T12 * (p->aigd + T3 * I->vgd_eff - T4 * I->vgd_eff * I->vgd_eff);
    if (T5 > OR_EXP_THRESHOLD) {
      T6 = OR_MAX_EXP;
      dT6_dVg = 0.0;
    } else if (T5 < -OR_EXP_THRESHOLD) {
      T6 = OR_MIN_EXP;
      dT6_dVg = 0.0;
    } else {
      T6 = exp(T5);
      dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * I->vgd_eff) * I->dvgd_eff_dvg;
    }
    I->Igd = T11 * T2 * T6;
    dIgd_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
    dIgd_dVd = -dIgd_dVg;
    I->gIgsg = dIgs_dVg;
    I->gIgss = dIgs_dVs;
    I->gIgdg = dIgd_dVg;
    I->gIgdd = dIgd_dVd;
  } else {
    I->Igcs = I->gIgcsg = I->gIgcsd = I->gIgcsb = 0.0;
    I->Igcd = I->gIgcdg = I->gIgcdd = I->gIgcdb = 0.0;
    I->Igs = I->gIgsg = I->gIgss = 0.0;
    I->Igd = I->gIgdg = I->gIgdd = 0.0;
  }
  if (M(igbMod)) {
    T0 = tmp * p->nigbacc;
    T1 = -Vgs_eff + Vbseff + Vfb;
    VxNVt = T1 / T0;
    if (VxNVt > OR_EXP_THRESHOLD) {
      Vaux = T1;
      dVaux_dVg = -dVgs_eff_dVg;
      dVaux_dVb = 1.0;
    } else if (VxNVt < -OR_EXP_THRESHOLD) {
      Vaux = T0 * log(1.0 + OR_MIN_EXP);
      dVaux_dVg = dVaux_dVb = 0.0;
    } else {
      ExpVxNVt = exp(VxNVt);
      Vaux = T0 * log(1.0 + ExpVxNVt);
      dVaux_dVb = ExpVxNVt / (1.0 + ExpVxNVt);
      dVaux_dVg = -dVaux_dVb * dVgs_eff_dVg;
    }
    T2 = (Vgs_eff - Vbseff) * Vaux;
    dT2_dVg = dVgs_eff_dVg * Vaux + (Vgs_eff - Vbseff) * dVaux_dVg;
    dT2_dVb = -Vaux + (Vgs_eff - Vbseff) * dVaux_dVb;
    T11 = 4.97232e-7 * p->weff * p->leff * p->ToxRatio;
    T12 = -7.45669e11 * toxe;
    T3 = p->aigbacc * p->cigbacc - p->bigbacc;
    T4 = p->bigbacc * p->cigbacc;
    T5 = T12 * (p->aigbacc + T3 * Voxacc - T4 * Voxacc * Voxacc);
    if (T5 > OR_EXP_THRESHOLD) {
      T6 = OR_MAX_EXP;
      dT6_dVg = dT6_dVb = 0.0;
    } else if (T5 < -OR_EXP_THRESHOLD) {
      T6 = OR_MIN_EXP;
      dT6_dVg = dT6_dVb = 0.0;
    } else {
      T6 = exp(T5);
      dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * Voxacc);
      dT6_dVb = dT6_dVg * dVoxacc_dVb;
      dT6_dVg *= dVoxacc_dVg;
    }
    Igbacc = T11 * T2 * T6;
    dIgbacc_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
    dIgbacc_dVb = T11 * (T2 * dT6_dVb + T6 * dT2_dVb);
    T0 = tmp * p->nigbinv;
    T1 = Voxdepinv - p->eigbinv;
    VxNVt = T1 / T0;
    if (VxNVt > OR_EXP_THRESHOLD) {
      Vaux = T1;
      dVaux_dVg = dVoxdepinv_dVg;
      dVaux_dVd = dVoxdepinv_dVd;
      dVaux_dVb = dVoxdepinv_dVb;
    } else if (VxNVt < -OR_EXP_THRESHOLD) {
      Vaux = T0 * log(1.0 + OR_MIN_EXP);
      dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
    } else {
      ExpVxNVt = exp(VxNVt);
      Vaux = T0 * log(1.0 + ExpVxNVt);
      dVaux_dVg = ExpVxNVt / (1.0 + ExpVxNVt);
      dVaux_dVd = dVaux_dVg * dVoxdepinv_dVd;
      dVaux_dVb = dVaux_dVg * dVoxdepinv_dVb;
      dVaux_dVg *= dVoxdepinv_dVg;
    }
    T2 = (Vgs_eff - Vbseff) * Vaux;
    dT2_dVg = dVgs_eff_dVg * Vaux + (Vgs_eff - Vbseff) * dVaux_dVg;
    dT2_dVd = (Vgs_eff - Vbseff) * dVaux_dVd;
    dT2_dVb = -Vaux + (Vgs_eff - Vbseff) * dVaux_dVb;
    T11 *= 0.75610;
    T12 *= 1.31724;
    T3 = p->aigbinv * p->cigbinv - p->bigbinv;
    T4 = p->bigbinv * p->cigbinv;
    T5 = T12 * (p->aigbinv + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
    if (T5 > OR_EXP_THRESHOLD) {
      T6 = OR_MAX_EXP;
      dT6_dVg = dT6_dVd = dT6_dVb = 0.0;
    } else if (T5 < -OR_EXP_THRESHOLD) {
      T6 = OR_MIN_EXP;
      dT6_dVg = dT6_dVd = dT6_dVb = 0.0;
    } else {
      T6 = exp(T5);
      dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * Voxdepinv);
      dT6_dVd = dT6_dVg * dVoxdepinv_dVd;
      dT6_dVb = dT6_dVg * dVoxdepinv_dVb;
      dT6_dVg *= dVoxdepinv_dVg;
    }
    Igbinv = T11 * T2 * T6;
    dIgbinv_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
    dIgbinv_dVd = T11 * (T2 * dT6_dVd + T6 * dT2_dVd);
    dIgbinv_dVb = T11 * (T2 * dT6_dVb + T6 * dT2_dVb);
    I->Igb = Igbinv + Igbacc;
    I->gIgbg = dIgbinv_dVg + dIgbacc_dVg;
    I->gIgbd = dIgbinv_dVd;
    I->gIgbb = (dIgbinv_dVb + dIgbacc_dVb) * dVbseff_dVb;
  } else {
    I->Igb = I->gIgbg = I->gIgbd = I->gIgbs = I->gIgbb = 0.0;
  }
  if (nf != 1.0) {
    I->cdrain *= nf;
    I->gds *= nf;
    I->gm *= nf;
    I->gmbs *= nf;
    I->IdovVds *= nf;
    I->gbbs *= nf;
    I->gbgs *= nf;
    I->gbds *= nf;
    I->csub *= nf;
    I->Igidl *= nf;
    I->ggidld *= nf;
    I->ggidlg *= nf;
    I->ggidlb *= nf;
    I->Igisl *= nf;
    I->ggisls *= nf;
    I->ggislg *= nf;
    I->ggislb *= nf;
    I->Igcs *= nf;
    I->gIgcsg *= nf;
    I->gIgcsd *= nf;
    I->gIgcsb *= nf;
    I->Igcd *= nf;
    I->gIgcdg *= nf;
    I->gIgcdd *= nf;
    I->gIgcdb *= nf;
    I->Igs *= nf;
    I->gIgsg *= nf;
    I->gIgss *= nf;
    I->Igd *= nf;
    I->gIgdg *= nf;
    I->gIgdd *= nf;
    I->Igb *= nf;
    I->gIgbg *= nf;
    I->gIgbd *= nf;
    I->gIgbb *= nf;
  }
  I->ggidls = -(I->ggidld + I->ggidlg + I->ggidlb);
  I->ggisld = -(I->ggisls + I->ggislg + I->ggislb);
  I->gIgbs = -(I->gIgbg + I->gIgbd + I->gIgbb);
  I->gIgcss = -(I->gIgcsg + I->gIgcsd + I->gIgcsb);
  I->gIgcds = -(I->gIgcdg + I->gIgcdd + I->gIgcdb);
  I->cd = I->cdrain;

  /* noise charge (:10674-10703): rewrites the Abulk and Vdseff members */
  if (M(tnoiMod) == 0) {
    I->Abulk = Abulk0 * p->abulkCVfactor;
    Vdsat = I->Vgsteff / I->Abulk;
    T0 = Vdsat - Vds - B4_DELTA_4;
    T1 = sqrt(T0 * T0 + 4.0 * B4_DELTA_4 * Vdsat);
    if (T0 >= 0.0) {
      I->Vdseff = Vdsat - 0.5 * (T0 + T1);
    } else {
      T3 = (B4_DELTA_4 + B4_DELTA_4) / (T1 - T0);
      T4 = 1.0 - T3;
      T5 = Vdsat * T3 / (T1 - T0);
      I->Vdseff = Vdsat * T4;
    }
    if (Vds == 0.0) I->Vdseff = 0.0;
    T0 = I->Abulk * I->Vdseff;
    T1 = 12.0 * (I->Vgsteff - 0.5 * T0 + 1.0e-20);
    T2 = I->Vdseff / T1;
    T3 = T0 * T2;
    I->qinv = Coxeff * p->weffCV * nf * p->leffCV * (I->Vgsteff - 0.5 * T0 + I->Abulk * T3);
  }

  /* intrinsic charges, capMod = 2 (:10707-11611) */
  if ((M(xpart) < 0) || (!I->ChargeComputationNeeded)) {
    I->qgate = I->qdrn = I->qsrc = I->qbulk = 0.0;
    I->cggb = I->cgsb = I->cgdb = 0.0;
    I->cdgb = I->cdsb = I->cddb = 0.0;
    I->cbgb = I->cbsb = I->cbdb = 0.0;
    I->csgb = I->cssb = I->csdb = 0.0;
    I->cgbb = I->csbb = I->cdbb = I->cbbb = 0.0;
    I->cqdb = I->cqsb = I->cqgb = I->cqbb = 0.0;
    I->gtau = 0.0;
  } else if (M(capMod) == 0) {
    /* capMod = 0 (:10718-11012) */
    if (Vbseff < 0.0) {
      Vbseff = Vbs;
      dVbseff_dVb = 1.0;
    } else {
      Vbseff = p->phi - Phis;
      dVbseff_dVb = -dPhis_dVb;
    }
    Vfb = p->vfbcv;
    Vth = Vfb + p->phi + p->k1ox * sqrtPhis;
    Vgst = Vgs_eff - Vth;
    dVth_dVb = p->k1ox * dsqrtPhis_dVb;
    dVgst_dVb = -dVth_dVb;
    dVgst_dVg = dVgs_eff_dVg;
    I->CoxWL = m->coxe * p->weffCV * p->leffCV * nf;
    Arg1 = Vgs_eff - Vbseff - Vfb;
    if (Arg1 <= 0.0) {
      I->qgate = I->CoxWL * Arg1;
      I->qbulk = -I->qgate;
      I->qdrn = 0.0;
      I->cggb = I->CoxWL * dVgs_eff_dVg;
      I->cgdb = 0.0;
      I->cgsb = I->CoxWL * (dVbseff_dVb - dVgs_eff_dVg);
      I->cdgb = 0.0;
      I->cddb = 0.0;
      I->cdsb = 0.0;
      I->cbgb = -I->CoxWL * dVgs_eff_dVg;
      I->cbdb = 0.0;
      I->cbsb = -I->cgsb;
    } else if (Vgst <= 0.0) {
      T1 = 0.5 * p->k1ox;
      T2 = sqrt(T1 * T1 + Arg1);
      I->qgate = I->CoxWL * p->k1ox * (T2 - T1);
      I->qbulk = -I->qgate;
      I->qdrn = 0.0;
      T0 = I->CoxWL * T1 / T2;
      I->cggb = T0 * dVgs_eff_dVg;
      I->cgdb = 0.0;
      I->cgsb = T0 * (dVbseff_dVb - dVgs_eff_dVg);
      I->cdgb = 0.0;
      I->cddb = 0.0;
      I->cdsb = 0.0;
      I->cbgb = -I->cggb;
      I->cbdb = 0.0;
      I->cbsb = -I->cgsb;
    } else {
      One_Third_CoxWL = I->CoxWL / 3.0;
      Two_Third_CoxWL = 2.0 * One_Third_CoxWL;
      AbulkCV = Abulk0 * p->abulkCVfactor;
      dAbulkCV_dVb = p->abulkCVfactor * dAbulk0_dVb;
      Vdsat = Vgst / AbulkCV;
      dVdsat_dVg = dVgs_eff_dVg / AbulkCV;
      dVdsat_dVb = -(Vdsat * dAbulkCV_dVb + dVth_dVb) / AbulkCV;
      if (M(xpart) > 0.5) {
        /* 0/100 partition */
        if (Vdsat <= Vds) {
          T1 = Vdsat / 3.0;
          I->qgate = I->CoxWL * (Vgs_eff - Vfb - p->phi - T1);
          T2 = -Two_Third_CoxWL * Vgst;
          I->qbulk = -(I->qgate + T2);
          I->qdrn = 0.0;
          I->cggb = One_Third_CoxWL * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
          T2 = -One_Third_CoxWL * dVdsat_dVb;
          I->cgsb = -(I->cggb + T2);
          I->cgdb = 0.0;
          I->cdgb = 0.0;
          I->cddb = 0.0;
          I->cdsb = 0.0;
          I->cbgb = -(I->cggb - Two_Third_CoxWL * dVgs_eff_dVg);
          T3 = -(T2 + Two_Third_CoxWL * dVth_dVb);
          I->cbsb = -(I->cbgb + T3);
          I->cbdb = 0.0;
        } else {
          Alphaz = Vgst / Vdsat;
          T1 = 2.0 * Vdsat - Vds;
          T2 = Vds / (3.0 * T1);
          T3 = T2 * Vds;
          T9 = 0.25 * I->CoxWL;
          T4 = T9 * Alphaz;
          T7 = 2.0 * Vds - T1 - 3.0 * T3;
          T8 = T3 - T1 - 2.0 * Vds;
          I->qgate = I->CoxWL * (Vgs_eff - Vfb - p->phi - 0.5 * (Vds - T3));
          T10 = T4 * T8;
          I->qdrn = T4 * T7;
          I->qbulk = -(I->qgate + I->qdrn + T10);
          T5 = T3 / T1;
          I->cggb = I->CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
          T11 = -I->CoxWL * T5 * dVdsat_dVb;
          I->cgdb = I->CoxWL * (T2 - 0.5 + 0.5 * T5);
          I->cgsb = -(I->cggb + T11 + I->cgdb);
          T6 = 1.0 / Vdsat;
          dAlphaz_dVg = T6 * (1.0 - Alphaz * dVdsat_dVg);
          dAlphaz_dVb = -T6 * (dVth_dVb + Alphaz * dVdsat_dVb);
          T7 = T9 * T7;
          T8 = T9 * T8;
          T9 = 2.0 * T4 * (1.0 - 3.0 * T5);
          I->cdgb = (T7 * dAlphaz_dVg - T9 * dVdsat_dVg) * dVgs_eff_dVg;
          T12 = T7 * dAlphaz_dVb - T9 * dVdsat_dVb;
          I->cddb = T4 * (3.0 - 6.0 * T2 - 3.0 * T5);
          I->cdsb = -(I->cdgb + T12 + I->cddb);
          T9 = 2.0 * T4 * (1.0 + T5);
          T10 = (T8 * dAlphaz_dVg - T9 * dVdsat_dVg) * dVgs_eff_dVg;
          T11 = T8 * dAlphaz_dVb - T9 * dVdsat_dVb;
          T12 = T4 * (2.0 * T2 + T5 - 1.0);
          T0 = -(T10 + T11 + T12);
          I->cbgb = -(I->cggb + I->cdgb + T10);
          I->cbdb = -(I->cgdb + I->cddb + T12);
          I->cbsb = -(I->cgsb + I->cdsb + T0);
        }
      } else if (M(xpart) < 0.5) {
        /* 40/60 partition */
        if (Vds >= Vdsat) {
          T1 = Vdsat / 3.0;
          I->qgate = I->CoxWL * (Vgs_eff - Vfb - p->phi - T1);
          T2 = -Two_Third_CoxWL * Vgst;
          I->qbulk = -(I->qgate + T2);
          I->qdrn = 0.4 * T2;
          I->cggb = One_Third_CoxWL * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
          T2 = -One_Third_CoxWL * dVdsat_dVb;
          I->cgsb = -(I->cggb + T2);
          I->cgdb = 0.0;
          T3 = 0.4 * Two_Third_CoxWL;
          I->cdgb = -T3 * dVgs_eff_dVg;
          I->cddb = 0.0;
          T4 = T3 * dVth_dVb;
          I->cdsb = -(T4 + I->cdgb);
          I->cbgb = -(I->cggb - Two_Third_CoxWL * dVgs_eff_dVg);
          T3 = -(T2 + Two_Third_CoxWL * dVth_dVb);
          I->cbsb = -(I->cbgb + T3);
          I->cbdb = 0.0;
        } else {
          Alphaz = Vgst / Vdsat;
          T1 = 2.0 * Vdsat - Vds;
          T2 = Vds / (3.0 * T1);
          T3 = T2 * Vds;
          T9 = 0.25 * I->CoxWL;
          T4 = T9 * Alphaz;
          I->qgate = I->CoxWL * (Vgs_eff - Vfb - p->phi - 0.5 * (Vds - T3));
          T5 = T3 / T1;
          I->cggb = I->CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
          tmp = -I->CoxWL * T5 * dVdsat_dVb;
          I->cgdb = I->CoxWL * (T2 - 0.5 + 0.5 * T5);
          I->cgsb = -(I->cggb + I->cgdb + tmp);
          T6 = 1.0 / Vdsat;
          dAlphaz_dVg = T6 * (1.0 - Alphaz * dVdsat_dVg);
          dAlphaz_dVb = -T6 * (dVth_dVb + Alphaz * dVdsat_dVb);
          T6 = 8.0 * Vdsat * Vdsat - 6.0 * Vdsat * Vds + 1.2 * Vds * Vds;
          T8 = T2 / T1;
          T7 = Vds - T1 - T8 * T6;
          I->qdrn = T4 * T7;
          T7 *= T9;
          tmp = T8 / T1;
          tmp1 = T4 * (2.0 - 4.0 * tmp * T6 + T8 * (16.0 * Vdsat - 6.0 * Vds));
          I->cdgb = (T7 * dAlphaz_dVg - tmp1 * dVdsat_dVg) * dVgs_eff_dVg;
          T10 = T7 * dAlphaz_dVb - tmp1 * dVdsat_dVb;
          I->cddb = T4 * (2.0 - (1.0 / (3.0 * T1 * T1) + 2.0 * tmp) * T6
                          + T8 * (6.0 * Vdsat - 2.4 * Vds));
          I->cdsb = -(I->cdgb + T10 + I->cddb);
          T7 = 2.0 * (T1 + T3);
          I->qbulk = -(I->qgate - T4 * T7);
          T7 *= T9;
          T0 = 4.0 * T4 * (1.0 - T5);
          T12 = (-T7 * dAlphaz_dVg - I->cdgb - T0 * dVdsat_dVg) * dVgs_eff_dVg;
          T11 = -T7 * dAlphaz_dVb - T10 - T0 * dVdsat_dVb;
          T10 = -4.0 * T4 * (T2 - 0.5 + 0.5 * T5) - I->cddb;
          tmp = -(T10 + T11 + T12);
          I->cbgb = -(I->cggb + I->cdgb + T12);
          I->cbdb = -(I->cgdb + I->cddb + T10);
          I->cbsb = -(I->cgsb + I->cdsb + tmp);
        }
      } else {
        /* 50/50 partition */
        if (Vds >= Vdsat) {
          T1 = Vdsat / 3.0;
          I->qgate = I->CoxWL * (Vgs_eff - Vfb - p->phi - T1);
          T2 = -Two_Third_CoxWL * Vgst;
          I->qbulk = -(I->qgate + T2);
          I->qdrn = 0.5 * T2;
          I->cggb = One_Third_CoxWL * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
          T2 = -One_Third_CoxWL * dVdsat_dVb;
          I->cgsb = -(I->cggb + T2);
          I->cgdb = 0.0;
          I->cdgb = -One_Third_CoxWL * dVgs_eff_dVg;
          I->cddb = 0.0;
          T4 = One_Third_CoxWL * dVth_dVb;
          I->cdsb = -(T4 + I->cdgb);
          I->cbgb = -(I->cggb - Two_Third_CoxWL * dVgs_eff_dVg);
          T3 = -(T2 + Two_Third_CoxWL * dVth_dVb);
          I->cbsb = -(I->cbgb + T3);
          I->cbdb = 0.0;
        } else {
          Alphaz = Vgst / Vdsat;
          T1 = 2.0 * Vdsat - Vds;
          T2 = Vds / (3.0 * T1);
          T3 = T2 * Vds;
          T9 = 0.25 * I->CoxWL;
          T4 = T9 * Alphaz;
          I->qgate = I->CoxWL * (Vgs_eff - Vfb - p->phi - 0.5 * (Vds - T3));
          T5 = T3 / T1;
          I->cggb = I->CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
          tmp = -I->CoxWL * T5 * dVdsat_dVb;
          I->cgdb = I->CoxWL * (T2 - 0.5 + 0.5 * T5);
          I->cgsb = -(I->cggb + I->cgdb + tmp);
          T6 = 1.0 / Vdsat;
          dAlphaz_dVg = T6 * (1.0 - Alphaz * dVdsat_dVg);
          dAlphaz_dVb = -T6 * (dVth_dVb + Alphaz * dVdsat_dVb);
          T7 = T1 + T3;
          I->qdrn = -T4 * T7;
          I->qbulk = -(I->qgate + I->qdrn + I->qdrn);
          T7 *= T9;
          T0 = T4 * (2.0 * T5 - 2.0);
          I->cdgb = (T0 * dVdsat_dVg - T7 * dAlphaz_dVg) * dVgs_eff_dVg;
          T12 = T0 * dVdsat_dVb - T7 * dAlphaz_dVb;
          I->cddb = T4 * (1.0 - 2.0 * T2 - T5);
          I->cdsb = -(I->cdgb + T12 + I->cddb);
          I->cbgb = -(I->cggb + 2.0 * I->cdgb);
          I->cbdb = -(I->cgdb + 2.0 * I->cddb);
          I->cbsb = -(I->cgsb + 2.0 * I->cdsb);
        }
      }
    }
  } else {
    if (Vbseff < 0.0) {
      VbseffCV = Vbseff;
      dVbseffCV_dVb = 1.0;
    } else {
      VbseffCV = p->phi - Phis;
      dVbseffCV_dVb = -dPhis_dVb;
    }
    I->CoxWL = m->coxe * p->weffCV * p->leffCV * nf;
    if (M(cvchargeMod) == 0) {
      noff = n * p->noff;
      dnoff_dVd = p->noff * dn_dVd;
      dnoff_dVb = p->noff * dn_dVb;
      T0 = Vtm * noff;
      voffcv = p->voffcv;
      VgstNVt = (Vgst - voffcv) / T0;
      if (VgstNVt > OR_EXP_THRESHOLD) {
        I->Vgsteff = Vgst - voffcv;
        dVgsteff_dVg = dVgs_eff_dVg;
        dVgsteff_dVd = -dVth_dVd;
        dVgsteff_dVb = -dVth_dVb;
      } else if (VgstNVt < -OR_EXP_THRESHOLD) {
        I->Vgsteff = T0 * log(1.0 + OR_MIN_EXP);
        dVgsteff_dVg = 0.0;
        dVgsteff_dVd = I->Vgsteff / noff;
        dVgsteff_dVb = dVgsteff_dVd * dnoff_dVb;
        dVgsteff_dVd *= dnoff_dVd;
      } else {
        ExpVgst = exp(VgstNVt);
        I->Vgsteff = T0 * log(1.0 + ExpVgst);
        dVgsteff_dVg = ExpVgst / (1.0 + ExpVgst);
        dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd)
                       + I->Vgsteff / noff * dnoff_dVd;
        dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb)
                       + I->Vgsteff / noff * dnoff_dVb;
        dVgsteff_dVg *= dVgs_eff_dVg;
      }
    } else {
      T0 = n * Vtm;
      T1 = p->mstarcv * Vgst;
      T2 = T1 / T0;
      if (T2 > OR_EXP_THRESHOLD) {
        T10 = T1;
        dT10_dVg = p->mstarcv * dVgs_eff_dVg;
        dT10_dVd = -dVth_dVd * p->mstarcv;
        dT10_dVb = -dVth_dVb * p->mstarcv;
      } else if (T2 < -OR_EXP_THRESHOLD) {
        T10 = Vtm * log(1.0 + OR_MIN_EXP);
        dT10_dVg = 0.0;
        dT10_dVd = T10 * dn_dVd;
        dT10_dVb = T10 * dn_dVb;
        T10 *= n;
      } else {
        ExpVgst = exp(T2);
        T3 = Vtm * log(1.0 + ExpVgst);
        T10 = n * T3;
        dT10_dVg = p->mstarcv * ExpVgst / (1.0 + ExpVgst);
        dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
        dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
        dT10_dVg *= dVgs_eff_dVg;
      }
      T1 = p->voffcbncv - (1.0 - p->mstarcv) * Vgst;
      T2 = T1 / T0;
      if (T2 < -OR_EXP_THRESHOLD) {
        T3 = m->coxe * OR_MIN_EXP / p->cdep0;
        T9 = p->mstarcv + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
      } else if (T2 > OR_EXP_THRESHOLD) {
        T3 = m->coxe * OR_MAX_EXP / p->cdep0;
        T9 = p->mstarcv + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
      } else {
        ExpVgst = exp(T2);
        T3 = m->coxe / p->cdep0;
        T4 = T3 * ExpVgst;
        T5 = T1 * T4 / T0;
        T9 = p->mstarcv + n * T4;
        dT9_dVg = T3 * (p->mstarcv - 1.0) * ExpVgst / Vtm;
        dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
        dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
        dT9_dVg *= dVgs_eff_dVg;
      }
      I->Vgsteff = T10 / T9;
      T11 = T9 * T9;
      dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
      dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
      dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;
    }
    if (M(capMod) == 1) {
    /* capMod = 1 (:11139-11319) */
    Vfb = I->vfbzb;
    V3 = Vfb - Vgs_eff + VbseffCV - B4_DELTA_3;
    if (Vfb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * B4_DELTA_3 * Vfb);
    else T0 = sqrt(V3 * V3 + 4.0 * B4_DELTA_3 * Vfb);
    T1 = 0.5 * (1.0 + V3 / T0);
    Vfbeff = Vfb - 0.5 * (V3 + T0);
    dVfbeff_dVg = T1 * dVgs_eff_dVg;
    dVfbeff_dVb = -T1 * dVbseffCV_dVb;
    Qac0 = I->CoxWL * (Vfbeff - Vfb);
    dQac0_dVg = I->CoxWL * dVfbeff_dVg;
    dQac0_dVb = I->CoxWL * dVfbeff_dVb;
    T0 = 0.5 * p->k1ox;
    T3 = Vgs_eff - Vfbeff - VbseffCV - I->Vgsteff;
    if (p->k1ox == 0.0) {
      T1 = 0.0;
      T2 = 0.0;
    } else if (T3 < 0.0) {
      T1 = T0 + T3 / p->k1ox;
      T2 = I->CoxWL;
    } else {
      T1 = sqrt(T0 * T0 + T3);
      T2 = I->CoxWL * T0 / T1;
    }
    Qsub0 = I->CoxWL * p->k1ox * (T1 - T0);
    dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg);
    dQsub0_dVd = -T2 * dVgsteff_dVd;
    dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb);
    AbulkCV = Abulk0 * p->abulkCVfactor;
    dAbulkCV_dVb = p->abulkCVfactor * dAbulk0_dVb;
    VdsatCV = I->Vgsteff / AbulkCV;
    T0 = VdsatCV - Vds - B4_DELTA_4;
    dT0_dVg = 1.0 / AbulkCV;
    dT0_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
    T1 = sqrt(T0 * T0 + 4.0 * B4_DELTA_4 * VdsatCV);
    dT1_dVg = (T0 + B4_DELTA_4 + B4_DELTA_4) / T1;
    dT1_dVd = -T0 / T1;
    dT1_dVb = dT1_dVg * dT0_dVb;
    dT1_dVg *= dT0_dVg;
    if (T0 >= 0.0) {
      VdseffCV = VdsatCV - 0.5 * (T0 + T1);
      dVdseffCV_dVg = 0.5 * (dT0_dVg - dT1_dVg);
      dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
      dVdseffCV_dVb = 0.5 * (dT0_dVb - dT1_dVb);
    } else {
      T3 = (B4_DELTA_4 + B4_DELTA_4) / (T1 - T0);
      T4 = 1.0 - T3;
      T5 = VdsatCV * T3 / (T1 - T0);
      VdseffCV = VdsatCV * T4;
      dVdseffCV_dVg = dT0_dVg * T4 + T5 * (dT1_dVg - dT0_dVg);
      dVdseffCV_dVd = T5 * (dT1_dVd + 1.0);
      dVdseffCV_dVb = dT0_dVb * (T4 - T5) + T5 * dT1_dVb;
    }
    if (Vds == 0.0) {
      VdseffCV = 0.0;
      dVdseffCV_dVg = 0.0;
      dVdseffCV_dVb = 0.0;
    }
    T0 = AbulkCV * VdseffCV;
    T1 = 12.0 * (I->Vgsteff - 0.5 * T0 + 1.0e-20);
    T2 = VdseffCV / T1;
    T3 = T0 * T2;
    T4 = (1.0 - 12.0 * T2 * T2 * AbulkCV);
    T5 = (6.0 * T0 * (4.0 * I->Vgsteff - T0) / (T1 * T1) - 0.5);
    T6 = 12.0 * T2 * T2 * I->Vgsteff;
    I->qgate = I->CoxWL * (I->Vgsteff - 0.5 * VdseffCV + T3);
    Cgg1 = I->CoxWL * (T4 + T5 * dVdseffCV_dVg);
    Cgd1 = I->CoxWL * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd;
    Cgb1 = I->CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteff_dVb;
    Cgg1 *= dVgsteff_dVg;
    T7 = 1.0 - AbulkCV;
    I->qbulk = I->CoxWL * T7 * (0.5 * VdseffCV - T3);
    T4 = -T7 * (T4 - 1.0);
    T5 = -T7 * T5;
    T6 = -(T7 * T6 + (0.5 * VdseffCV - T3));
    Cbg1 = I->CoxWL * (T4 + T5 * dVdseffCV_dVg);
    Cbd1 = I->CoxWL * T5 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd;
    Cbb1 = I->CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb;
    Cbg1 *= dVgsteff_dVg;
    if (M(xpart) > 0.5) {
      T1 = T1 + T1;
      I->qsrc = -I->CoxWL * (0.5 * I->Vgsteff + 0.25 * T0 - T0 * T0 / T1);
      T7 = (4.0 * I->Vgsteff - T0) / (T1 * T1);
      T4 = -(0.5 + 24.0 * T0 * T0 / (T1 * T1));
      T5 = -(0.25 * AbulkCV - 12.0 * AbulkCV * T0 * T7);
      T6 = -(0.25 * VdseffCV - 12.0 * T0 * VdseffCV * T7);
      Csg = I->CoxWL * (T4 + T5 * dVdseffCV_dVg);
      Csd = I->CoxWL * T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd;
      Csb = I->CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb;
      Csg *= dVgsteff_dVg;
    } else if (M(xpart) < 0.5) {
      T1 = T1 / 12.0;
      T2 = 0.5 * I->CoxWL / (T1 * T1);
      T3 = I->Vgsteff * (2.0 * T0 * T0 / 3.0 + I->Vgsteff * (I->Vgsteff - 4.0 * T0 / 3.0))
           - 2.0 * T0 * T0 * T0 / 15.0;
      I->qsrc = -T2 * T3;
      T7 = 4.0 / 3.0 * I->Vgsteff * (I->Vgsteff - T0) + 0.4 * T0 * T0;
      T4 = -2.0 * I->qsrc / T1 - T2 * (I->Vgsteff * (3.0 * I->Vgsteff - 8.0 * T0 / 3.0)
                                       + 2.0 * T0 * T0 / 3.0);
      T5 = (I->qsrc / T1 + T2 * T7) * AbulkCV;
      T6 = (I->qsrc / T1 * VdseffCV + T2 * T7 * VdseffCV);
      Csg = (T4 + T5 * dVdseffCV_dVg);
      Csd = T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd;
      Csb = (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb;
      Csg *= dVgsteff_dVg;
    } else {
      I->qsrc = -0.5 * (I->qgate + I->qbulk);
      Csg = -0.5 * (Cgg1 + Cbg1);
      Csb = -0.5 * (Cgb1 + Cbb1);
      Csd = -0.5 * (Cgd1 + Cbd1);
    }
    I->qgate += Qac0 + Qsub0;
    I->qbulk -= (Qac0 + Qsub0);
    I->qdrn = -(I->qgate + I->qbulk + I->qsrc);
    Cgg = dQac0_dVg + dQsub0_dVg + Cgg1;
    Cgd = dQsub0_dVd + Cgd1;
    Cgb = dQac0_dVb + dQsub0_dVb + Cgb1;
    Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
    Cbd = Cbd1 - dQsub0_dVd;
    Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
    Cgb *= dVbseff_dVb;
    Cbb *= dVbseff_dVb;
    Csb *= dVbseff_dVb;
    I->cggb = Cgg;
    I->cgsb = -(Cgg + Cgd + Cgb);
    I->cgdb = Cgd;
    I->cdgb = -(Cgg + Cbg + Csg);
    I->cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
    I->cddb = -(Cgd + Cbd + Csd);
    I->cbgb = Cbg;
    I->cbsb = -(Cbg + Cbd + Cbb);
    I->cbdb = Cbd;
    } else {
    /* capMod = 2 (:11321-11602) */
    V3 = I->vfbzb - Vgs_eff + VbseffCV - B4_DELTA_3;
    if (I->vfbzb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * B4_DELTA_3 * I->vfbzb);
    else T0 = sqrt(V3 * V3 + 4.0 * B4_DELTA_3 * I->vfbzb);
    T1 = 0.5 * (1.0 + V3 / T0);
    Vfbeff = I->vfbzb - 0.5 * (V3 + T0);
    dVfbeff_dVg = T1 * dVgs_eff_dVg;
    dVfbeff_dVb = -T1 * dVbseffCV_dVb;
    Cox = m->coxp;
    Tox = 1.0e8 * M(toxp);
    T0 = (Vgs_eff - VbseffCV - I->vfbzb) / Tox;
    dT0_dVg = dVgs_eff_dVg / Tox;
    dT0_dVb = -dVbseffCV_dVb / Tox;
    tmp = T0 * p->acde;
    if ((-OR_EXP_THRESHOLD < tmp) && (tmp < OR_EXP_THRESHOLD)) {
      Tcen = p->ldeb * exp(tmp);
      dTcen_dVg = p->acde * Tcen;
      dTcen_dVb = dTcen_dVg * dT0_dVb;
      dTcen_dVg *= dT0_dVg;
    } else if (tmp <= -OR_EXP_THRESHOLD) {
      Tcen = p->ldeb * OR_MIN_EXP;
      dTcen_dVg = dTcen_dVb = 0.0;
    } else {
      Tcen = p->ldeb * OR_MAX_EXP;
      dTcen_dVg = dTcen_dVb = 0.0;
    }
    LINK = 1.0e-3 * M(toxp);
    V3 = p->ldeb - Tcen - LINK;
    V4 = sqrt(V3 * V3 + 4.0 * LINK * p->ldeb);
    Tcen = p->ldeb - 0.5 * (V3 + V4);
    T1 = 0.5 * (1.0 + V3 / V4);
    dTcen_dVg *= T1;
    dTcen_dVb *= T1;
    Ccen = epssub / Tcen;
    T2 = Cox / (Cox + Ccen);
    Coxeff = T2 * Ccen;
    T3 = -Ccen / Tcen;
    dCoxeff_dVg = T2 * T2 * T3;
    dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
    dCoxeff_dVg *= dTcen_dVg;
    CoxWLcen = I->CoxWL * Coxeff / m->coxe;
    Qac0 = CoxWLcen * (Vfbeff - I->vfbzb);
    QovCox = Qac0 / Coxeff;
    dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
    dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
    T0 = 0.5 * p->k1ox;
    T3 = Vgs_eff - Vfbeff - VbseffCV - I->Vgsteff;
    if (p->k1ox == 0.0) {
      T1 = 0.0;
      T2 = 0.0;
    } else if (T3 < 0.0) {
      T1 = T0 + T3 / p->k1ox;
      T2 = CoxWLcen;
    } else {
      T1 = sqrt(T0 * T0 + T3);
      T2 = CoxWLcen * T0 / T1;
    }
    Qsub0 = CoxWLcen * p->k1ox * (T1 - T0);
    QovCox = Qsub0 / Coxeff;
    dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg) + QovCox * dCoxeff_dVg;
    dQsub0_dVd = -T2 * dVgsteff_dVd;
    dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb) + QovCox * dCoxeff_dVb;
    if (p->k1ox <= 0.0) {
      Denomi = 0.25 * p->moin * Vtm;
      T0 = 0.5 * p->sqrtPhi;
    } else {
      Denomi = p->moin * Vtm * p->k1ox * p->k1ox;
      T0 = p->k1ox * p->sqrtPhi;
    }
    T1 = 2.0 * T0 + I->Vgsteff;
    DeltaPhi = Vtm * log(1.0 + T1 * I->Vgsteff / Denomi);
    dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (Denomi + T1 * I->Vgsteff);
    T0 = I->Vgsteff - DeltaPhi - 0.001;
    dT0_dVg = 1.0 - dDeltaPhi_dVg;
    T1 = sqrt(T0 * T0 + I->Vgsteff * 0.004);
    VgDP = 0.5 * (T0 + T1);
    dVgDP_dVg = 0.5 * (dT0_dVg + (T0 * dT0_dVg + 0.002) / T1);
    Tox += Tox;
    T0 = (I->Vgsteff + I->vtfbphi2) / Tox;
    tmp = exp(M(bdos) * 0.7 * log(T0));
    T1 = 1.0 + tmp;
    T2 = M(bdos) * 0.7 * tmp / (T0 * Tox);
    Tcen = M(ados) * 1.9e-9 / T1;
    dTcen_dVg = -Tcen * T2 / T1;
    dTcen_dVd = dTcen_dVg * dVgsteff_dVd;
    dTcen_dVb = dTcen_dVg * dVgsteff_dVb;
    dTcen_dVg *= dVgsteff_dVg;
    Ccen = epssub / Tcen;
    T0 = Cox / (Cox + Ccen);
    Coxeff = T0 * Ccen;
    T1 = -Ccen / Tcen;
    dCoxeff_dVg = T0 * T0 * T1;
    dCoxeff_dVd = dCoxeff_dVg * dTcen_dVd;
    dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
    dCoxeff_dVg *= dTcen_dVg;
    CoxWLcen = I->CoxWL * Coxeff / m->coxe;
    AbulkCV = Abulk0 * p->abulkCVfactor;
    dAbulkCV_dVb = p->abulkCVfactor * dAbulk0_dVb;
    VdsatCV = VgDP / AbulkCV;
    T0 = VdsatCV - Vds - B4_DELTA_4;
    dT0_dVg = dVgDP_dVg / AbulkCV;
    dT0_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
    T1 = sqrt(T0 * T0 + 4.0 * B4_DELTA_4 * VdsatCV);
    dT1_dVg = (T0 + B4_DELTA_4 + B4_DELTA_4) / T1;
    dT1_dVd = -T0 / T1;
    dT1_dVb = dT1_dVg * dT0_dVb;
    dT1_dVg *= dT0_dVg;
    if (T0 >= 0.0) {
      VdseffCV = VdsatCV - 0.5 * (T0 + T1);
      dVdseffCV_dVg = 0.5 * (dT0_dVg - dT1_dVg);
      dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
      dVdseffCV_dVb = 0.5 * (dT0_dVb - dT1_dVb);
    } else {
      T3 = (B4_DELTA_4 + B4_DELTA_4) / (T1 - T0);
      T4 = 1.0 - T3;
      T5 = VdsatCV * T3 / (T1 - T0);
      VdseffCV = VdsatCV * T4;
      dVdseffCV_dVg = dT0_dVg * T4 + T5 * (dT1_dVg - dT0_dVg);
      dVdseffCV_dVd = T5 * (dT1_dVd + 1.0);
      dVdseffCV_dVb = dT0_dVb * (T4 - T5) + T5 * dT1_dVb;
    }
    if (Vds == 0.0) {
      VdseffCV = 0.0;
      dVdseffCV_dVg = 0.0;
      dVdseffCV_dVb = 0.0;
    }
    T0 = AbulkCV * VdseffCV;
    T1 = VgDP;
    T2 = 12.0 * (T1 - 0.5 * T0 + 1.0e-20);
    T3 = T0 / T2;
    T4 = 1.0 - 12.0 * T3 * T3;
    T5 = AbulkCV * (6.0 * T0 * (4.0 * T1 - T0) / (T2 * T2) - 0.5);
    T6 = T5 * VdseffCV / AbulkCV;
    I->qgate = CoxWLcen * (T1 - T0 * (0.5 - T3));
    QovCox = I->qgate / Coxeff;
    Cgg1 = CoxWLcen * (T4 * dVgDP_dVg + T5 * dVdseffCV_dVg);
    Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
    Cgb1 = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
    Cgg1 = Cgg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
    T7 = 1.0 - AbulkCV;
    T8 = T2 * T2;
    T9 = 12.0 * T7 * T0 * T0 / (T8 * AbulkCV);
    T10 = T9 * dVgDP_dVg;
    T11 = -T7 * T5 / AbulkCV;
    T12 = -(T9 * T1 / AbulkCV + VdseffCV * (0.5 - T0 / T2));
    I->qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
    QovCox = I->qbulk / Coxeff;
    Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
    Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
    Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
    Cbg1 = Cbg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
    if (M(xpart) > 0.5) {
      I->qsrc = -CoxWLcen * (T1 / 2.0 + T0 / 4.0 - 0.5 * T0 * T0 / T2);
      QovCox = I->qsrc / Coxeff;
      T2 += T2;
      T3 = T2 * T2;
      T7 = -(0.25 - 12.0 * T0 * (4.0 * T1 - T0) / T3);
      T4 = -(0.5 + 24.0 * T0 * T0 / T3) * dVgDP_dVg;
      T5 = T7 * AbulkCV;
      T6 = T7 * VdseffCV;
      Csg = CoxWLcen * (T4 + T5 * dVdseffCV_dVg);
      Csd = CoxWLcen * T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd + QovCox * dCoxeff_dVd;
      Csb = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb + QovCox * dCoxeff_dVb;
      Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
    } else if (M(xpart) < 0.5) {
      T2 = T2 / 12.0;
      T3 = 0.5 * CoxWLcen / (T2 * T2);
      T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
      I->qsrc = -T3 * T4;
      QovCox = I->qsrc / Coxeff;
      T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
      T5 = -2.0 * I->qsrc / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
      T6 = AbulkCV * (I->qsrc / T2 + T3 * T8);
      T7 = T6 * VdseffCV / AbulkCV;
      Csg = T5 * dVgDP_dVg + T6 * dVdseffCV_dVg;
      Csd = Csg * dVgsteff_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
      Csb = Csg * dVgsteff_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
      Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
    } else {
      I->qsrc = -0.5 * I->qgate;
      Csg = -0.5 * Cgg1;
      Csd = -0.5 * Cgd1;
      Csb = -0.5 * Cgb1;
    }
    I->qgate += Qac0 + Qsub0 - I->qbulk;
    I->qbulk -= (Qac0 + Qsub0);
    I->qdrn = -(I->qgate + I->qbulk + I->qsrc);
    Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
    Cbd = Cbd1 - dQsub0_dVd;
    Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
    Cgg = Cgg1 - Cbg;
    Cgd = Cgd1 - Cbd;
    Cgb = Cgb1 - Cbb;
    Cgb *= dVbseff_dVb;
    Cbb *= dVbseff_dVb;
    Csb *= dVbseff_dVb;
    I->cggb = Cgg;
    I->cgsb = -(Cgg + Cgd + Cgb);
    I->cgdb = Cgd;
    I->cdgb = -(Cgg + Cbg + Csg);
    I->cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
    I->cddb = -(Cgd + Cbd + Csd);
    I->cbgb = Cbg;
    I->cbsb = -(Cbg + Cbd + Cbb);
    I->cbdb = Cbd;
    }
  }
  I->csgb = -I->cggb - I->cdgb - I->cbgb;
  I->csdb = -I->cgdb - I->cddb - I->cbdb;
  I->cssb = -I->cgsb - I->cdsb - I->cbsb;
  I->cgbb = -I->cgdb - I->cggb - I->cgsb;
  I->cdbb = -I->cddb - I->cdgb - I->cdsb;
  I->cbbb = -I->cbgb - I->cbdb - I->cbsb;
  I->csbb = -I->cgbb - I->cdbb - I->cbbb;

  /* NQS (:11626-11660) */
  if ((I->itrnqsMod) || (I->iacnqsMod)) {
    I->qchqs = I->qcheq = -(I->qbulk + I->qgate);
    I->cqgb = -(I->cggb + I->cbgb);
    I->cqdb = -(I->cgdb + I->cbdb);
    I->cqsb = -(I->cgsb + I->cbsb);
    I->cqbb = -(I->cqgb + I->cqdb + I->cqsb);
    I->CoxWL = m->coxe * p->weffCV * nf * p->leffCV;
    T1 = I->gcrg / I->CoxWL;
    I->gtau = T1 * I->ScalingFactor;
    if (I->iacnqsMod) I->taunet = 1.0 / T1;
  }

  /* junction charges (:11662-11810) */
  if (I->ChargeComputationNeeded) {
    czbd = m->DunitAreaTempJctCap * I->Adeff;
    czbs = m->SunitAreaTempJctCap * I->Aseff;
    czbdsw = m->DunitLengthSidewallTempJctCap * I->Pdeff;
    czbdswg = m->DunitLengthGateSidewallTempJctCap * p->weffCJ * nf;
    czbssw = m->SunitLengthSidewallTempJctCap * I->Pseff;
    czbsswg = m->SunitLengthGateSidewallTempJctCap * p->weffCJ * nf;
    MJS = M(SbulkJctBotGradingCoeff);
    MJSWS = M(SbulkJctSideGradingCoeff);
    MJSWGS = M(SbulkJctGateSideGradingCoeff);
    MJD = M(DbulkJctBotGradingCoeff);
    MJSWD = M(DbulkJctSideGradingCoeff);
    MJSWGD = M(DbulkJctGateSideGradingCoeff);
    if (I->vbs_jct == 0.0) {
      I->qbs = 0.0;
      I->capbs = czbs + czbssw + czbsswg;
    } else if (I->vbs_jct < 0.0) {
      if (czbs > 0.0) {
        arg = 1.0 - I->vbs_jct / m->PhiBS;
        if (MJS == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJS * log(arg));
        I->qbs = m->PhiBS * czbs * (1.0 - arg * sarg) / (1.0 - MJS);
        I->capbs = czbs * sarg;
      } else {
        I->qbs = 0.0;
        I->capbs = 0.0;
      }
      if (czbssw > 0.0) {
        arg = 1.0 - I->vbs_jct / m->PhiBSWS;
        if (MJSWS == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWS * log(arg));
        I->qbs += m->PhiBSWS * czbssw * (1.0 - arg * sarg) / (1.0 - MJSWS);
        I->capbs += czbssw * sarg;
      }
      if (czbsswg > 0.0) {
        arg = 1.0 - I->vbs_jct / m->PhiBSWGS;
        if (MJSWGS == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWGS * log(arg));
        I->qbs += m->PhiBSWGS * czbsswg * (1.0 - arg * sarg) / (1.0 - MJSWGS);
        I->capbs += czbsswg * sarg;
      }
    } else {
      T0 = czbs + czbssw + czbsswg;
      T1 = I->vbs_jct * (czbs * MJS / m->PhiBS + czbssw * MJSWS / m->PhiBSWS + czbsswg * MJSWGS / m->PhiBSWGS);
      I->qbs = I->vbs_jct * (T0 + 0.5 * T1);
      I->capbs = T0 + T1;
    }
    if (I->vbd_jct == 0.0) {
      I->qbd = 0.0;
      I->capbd = czbd + czbdsw + czbdswg;
    } else if (I->vbd_jct < 0.0) {
      if (czbd > 0.0) {
        arg = 1.0 - I->vbd_jct / m->PhiBD;
        if (MJD == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJD * log(arg));
        I->qbd = m->PhiBD * czbd * (1.0 - arg * sarg) / (1.0 - MJD);
        I->capbd = czbd * sarg;
      } else {
        I->qbd = 0.0;
        I->capbd = 0.0;
      }
      if (czbdsw > 0.0) {
        arg = 1.0 - I->vbd_jct / m->PhiBSWD;
        if (MJSWD == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWD * log(arg));
        I->qbd += m->PhiBSWD * czbdsw * (1.0 - arg * sarg) / (1.0 - MJSWD);
        I->capbd += czbdsw * sarg;
      }
      if (czbdswg > 0.0) {
        arg = 1.0 - I->vbd_jct / m->PhiBSWGD;
        if (MJSWGD == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWGD * log(arg));
        I->qbd += m->PhiBSWGD * czbdswg * (1.0 - arg * sarg) / (1.0 - MJSWGD);
        I->capbd += czbdswg * sarg;
      }
    } else {
      T0 = czbd + czbdsw + czbdswg;
      T1 = I->vbd_jct * (czbd * MJD / m->PhiBD + czbdsw * MJSWD / m->PhiBSWD + czbdswg * MJSWGD / m->PhiBSWGD);
      I->qbd = I->vbd_jct * (T0 + 0.5 * T1);
      I->capbd = T0 + T1;
    }
  }

  /* gate resistor currents and overlap capacitances (:11812-11883) */
  if (I->irgateMod == 3) {
    vgdx = I->vgmd;
    vgsx = I->vgms;
  } else {
    vgdx = I->vgd;
    vgsx = I->vgs;
  }
  I->Igate = I->IgateMid = 0.0;
  I->Igate_Jdxp = I->IgateMid_Jdxp = 0.0;
  if (I->irgateMod == 1) {
    I->Igate = I->grgeltd * (I->Vgegp);
  } else if (I->irgateMod == 2) {
    I->Igate = (I->gcrg) * (I->Vgegp);
  } else if (I->irgateMod == 3) {
    I->Igate = I->grgeltd * (I->Vgegm);
    I->IgateMid = I->gcrg * (I->Vgmgp);
  }
  if (M(capMod) == 0) {
    I->cgdo = p->cgdo;
    I->qgdo = p->cgdo * vgdx;
    I->cgso = p->cgso;
    I->qgso = p->cgso * vgsx;
  } else {
  T0 = vgdx + B4_DELTA_1;
  T1 = sqrt(T0 * T0 + 4.0 * B4_DELTA_1);
  T2 = 0.5 * (T0 - T1);
  T3 = p->weffCV * p->cgdl;
  T4 = sqrt(1.0 - 4.0 * T2 / p->ckappad);
  I->cgdo = p->cgdo + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
  I->qgdo = (p->cgdo + T3) * vgdx - T3 * (T2 + 0.5 * p->ckappad * (T4 - 1.0));
  T0 = vgsx + B4_DELTA_1;
  T1 = sqrt(T0 * T0 + 4.0 * B4_DELTA_1);
  T2 = 0.5 * (T0 - T1);
  T3 = p->weffCV * p->cgsl;
  T4 = sqrt(1.0 - 4.0 * T2 / p->ckappas);
  I->cgso = p->cgso + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
  I->qgso = (p->cgso + T3) * vgsx - T3 * (T2 + 0.5 * p->ckappas * (T4 - 1.0));
  }
  if (nf != 1.0) {
    I->cgdo *= nf;
    I->cgso *= nf;
    I->qgdo *= nf;
    I->qgso *= nf;
  }
  b4_setup_capacitors_old(I, ss);
  b4_setup_capacitors_new(I);
  if (M(rdsMod) == 1) {
    I->Idrain = I->gdtot * I->Vddp;
    I->Isource = I->gstot * I->Vssp;
  } else {
    I->Idrain = I->drainConductance * I->Vddp;
    I->Isource = I->sourceConductance * I->Vssp;
  }
  /* body network currents: the reference tests the model's rbodyMod here */
  if (M(rbodyMod) != 0) {
    I->Idbb = I->grbdb * I->Vdbb;
    I->Idbbp = I->grbpd * I->Vdbbp;
    I->Isbb = I->grbsb * I->Vsbb;
    I->Isbbp = I->grbps * I->Vsbbp;
    I->Ibpb = I->grbpb * I->Vbpb;
  }
  (void)Vgs; (void)dTcen_dVd;
  return 1;
}
