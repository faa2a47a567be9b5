/* or_bsim4.h -- oracle restatement of the Xyce BSIM4 (level 54) device:
 * data layout (test infrastructure only).
 * Reference: src/DeviceModelPKG/OpenModels/include/N_DEV_MOSFET_B4.h
 *   SizeDependParam :88-321, Instance :322-1149, Model :1151-2138.
 * Field names follow the reference so that each restated statement reads
 * like the line it restates. */
#ifndef OR_BSIM4_H
#define OR_BSIM4_H
#include "or_internal.h"

#define B4_EPS0      (8.854214871e-12)      /* N_DEV_Const.h CONSTperm0 */
#define B4_EPSSI     (1.03594e-10)
#define B4_KboQ      (8.617087e-5)          /* N_DEV_MOSFET_B4.C:106 */
#define B4_vt0       (OR_BOLTZ * (27.0 + OR_CtoK) / OR_Q)
#define B4_MM        3
#define B4_DELTA     (1.0E-9)
#define B4_DELTA_1   (0.02)
#define B4_DELTA_3   (0.02)
#define B4_DELTA_4   (0.02)
#define B4_Charge_q  (1.60219e-19)

typedef struct { char dummy; } b4_unused;

/* ---- model card: every table parameter as a named field ------------- */
typedef struct {
#define B4M(NAME, mem, def) double mem;
#include "../include/xa_bsim4_params.def"
#undef B4M
} b4params;

typedef struct {
#define B4M(NAME, mem, def) char mem;
#include "../include/xa_bsim4_params.def"
#undef B4M
} b4given;

typedef struct b4size_s b4size;

typedef struct {
  b4params p;           /* M(x) */
  b4given g;            /* MG(x) */
  int dtype;            /* +1 NMOS, -1 PMOS */
  /* derived, temperature dependent (written by every instance's
     updateTemperature; the last one wins, as in the reference) */
  double coxe, coxp, vcrit, factor1, vtm0, vtm, Eg0;
  double SjctTempSatCurDensity, SjctSidewallTempSatCurDensity, SjctGateSidewallTempSatCurDensity;
  double DjctTempSatCurDensity, DjctSidewallTempSatCurDensity, DjctGateSidewallTempSatCurDensity;
  double SunitAreaTempJctCap, DunitAreaTempJctCap;
  double SunitLengthSidewallTempJctCap, DunitLengthSidewallTempJctCap;
  double SunitLengthGateSidewallTempJctCap, DunitLengthGateSidewallTempJctCap;
  double PhiBS, PhiBD, PhiBSWS, PhiBSWD, PhiBSWGS, PhiBSWGD;
  double njtsstemp, njtsswstemp, njtsswgstemp, njtsdtemp, njtsswdtemp, njtsswgdtemp;
  /* size-dependent parameter cache keyed by (L, W, NF) (N_DEV_MOSFET_B4.C:6018-6049) */
  b4size** sizes;
  int nsizes, capsizes;
} b4model;

/* ---- size-dependent parameters (SizeDependParam) ---------------------- */
struct b4size_s {
  double Length, Width, NFinger, referenceTemperature;
#define B4BIN(sm, base, l, w, p) double sm;
#include "../include/xa_bsim4_binning.def"
#undef B4BIN
  double dl, dlc, dw, dwc, dwj, leff, weff, leffCV, weffCV, weffCJ;
  double abulkCVfactor, vsattemp, rds0, rdswmin, rd0, rdwmin, rs0, rswmin;
  double u0temp, tfactor, lc, cgdo, cgso, cgbo, phi, sqrtPhi, phis3, Xdep0, sqrtXdep0;
  double litl, vbi, vfbsd, cdep0, ToxRatio, ToxRatioEdge, Aechvb, Bechvb;
  double AechvbEdgeS, AechvbEdgeD, BechvbEdge, mstar, mstarcv, voffcbn, voffcbncv;
  double ldeb, k1ox, theta0vb0, thetaRout, vfbzbfactor, ku0, kvth0, ku0temp;
  double inv_od_ref, rho_ref;
};

/* ---- Jacobian pairs -------------------------------------------------------
 * Every (row node, column node) pair loadDAEMatrices addresses
 * (N_DEV_MOSFET_B4.C:16577-17268), named as the reference's matrix pointers:
 * GE gate (external), GP gate prime, GM gate mid, D/S/B external,
 * DP/SP/BP primes, DB/SB drain/source body, Q the NQS charge, IBS/IDS/IGS the
 * initial-condition branch currents. */
#define B4_PAIRS(X) \
  X(GEge, GateExt, GateExt) X(GEgp, GateExt, GatePrime) X(GEdp, GateExt, DrainPrime) \
  X(GEsp, GateExt, SourcePrime) X(GEbp, GateExt, BodyPrime) X(GEgm, GateExt, GateMid) \
  X(GPge, GatePrime, GateExt) X(GPgp, GatePrime, GatePrime) X(GPdp, GatePrime, DrainPrime) \
  X(GPsp, GatePrime, SourcePrime) X(GPbp, GatePrime, BodyPrime) X(GPgm, GatePrime, GateMid) \
  X(GPq, GatePrime, Charge) \
  X(GMge, GateMid, GateExt) X(GMgm, GateMid, GateMid) X(GMdp, GateMid, DrainPrime) \
  X(GMgp, GateMid, GatePrime) X(GMsp, GateMid, SourcePrime) X(GMbp, GateMid, BodyPrime) \
  X(Dgp, Drain, GatePrime) X(Dsp, Drain, SourcePrime) X(Dbp, Drain, BodyPrime) \
  X(Ddp, Drain, DrainPrime) X(Dd, Drain, Drain) X(Dids, Drain, Ids) \
  X(Sdp, Source, DrainPrime) X(Sgp, Source, GatePrime) X(Sbp, Source, BodyPrime) \
  X(Ssp, Source, SourcePrime) X(Ss, Source, Source) X(Sibs, Source, Ibs) X(Sids, Source, Ids) \
  X(Sigs, Source, Igs) \
  X(DPdp, DrainPrime, DrainPrime) X(DPd, DrainPrime, Drain) X(DPgp, DrainPrime, GatePrime) \
  X(DPsp, DrainPrime, SourcePrime) X(DPbp, DrainPrime, BodyPrime) X(DPdb, DrainPrime, DrainBody) \
  X(DPq, DrainPrime, Charge) X(DPgm, DrainPrime, GateMid) \
  X(SPdp, SourcePrime, DrainPrime) X(SPgp, SourcePrime, GatePrime) \
  X(SPsp, SourcePrime, SourcePrime) X(SPs, SourcePrime, Source) X(SPbp, SourcePrime, BodyPrime) \
  X(SPsb, SourcePrime, SourceBody) X(SPq, SourcePrime, Charge) X(SPgm, SourcePrime, GateMid) \
  X(BPdp, BodyPrime, DrainPrime) X(BPgp, BodyPrime, GatePrime) X(BPsp, BodyPrime, SourcePrime) \
  X(BPbp, BodyPrime, BodyPrime) X(BPdb, BodyPrime, DrainBody) X(BPb, BodyPrime, Body) \
  X(BPsb, BodyPrime, SourceBody) X(BPgm, BodyPrime, GateMid) \
  X(DBdp, DrainBody, DrainPrime) X(DBdb, DrainBody, DrainBody) X(DBbp, DrainBody, BodyPrime) \
  X(DBb, DrainBody, Body) \
  X(SBsp, SourceBody, SourcePrime) X(SBbp, SourceBody, BodyPrime) X(SBb, SourceBody, Body) \
  X(SBsb, SourceBody, SourceBody) \
  X(Bdb, Body, DrainBody) X(Bbp, Body, BodyPrime) X(Bsb, Body, SourceBody) X(Bb, Body, Body) \
  X(Bibs, Body, Ibs) \
  X(Qq, Charge, Charge) X(Qgp, Charge, GatePrime) X(Qdp, Charge, DrainPrime) \
  X(Qsp, Charge, SourcePrime) X(Qbp, Charge, BodyPrime) \
  X(IBSb, Ibs, Body) X(IBSs, Ibs, Source) X(IBSibs, Ibs, Ibs) \
  X(IDSd, Ids, Drain) X(IDSs, Ids, Source) X(IDSids, Ids, Ids) \
  X(GEigs, GateExt, Igs) X(IGSg, Igs, GateExt) X(IGSs, Igs, Source) X(IGSigs, Igs, Igs)
enum {
#define B4_PAIR_ENUM(name, r, c) B4_##name,
  B4_PAIRS(B4_PAIR_ENUM)
#undef B4_PAIR_ENUM
  B4_NPAIR
};

/* ---- instance ----------------------------------------------------------- */
typedef struct {
  b4model* model;
  b4size* paramPtr;
  /* instance parameters (B4I) and given flags */
#define B4I(NAME, mem, def) double mem;
#include "../include/xa_bsim4_params.def"
#undef B4I
#define B4I(NAME, mem, def) char mem##_given;
#include "../include/xa_bsim4_params.def"
#undef B4I
  int irbodyMod, irgateMod, igeoMod, irgeoMod, itrnqsMod, iacnqsMod, imin;
  int drainMOSFET_B4Exists, sourceMOSFET_B4Exists;
  int loadLeadCurrent;
  /* temperature / geometry dependent */
  double u0temp, vsattemp, vth0, vfb, eta0, k2, k2ox, vtfbphi1, vtfbphi2, vbsc, vfbzb;
  double cgso, cgdo, grbdb, grbpb, grbps, grbsb, grbpd, grgeltd;
  double DMCGeff, DMCIeff, DMDGeff, Pseff, Pdeff, Aseff, Adeff;
  double sourceConductance, drainConductance;
  double XExpBVS, XExpBVD, vjsmFwd, vjsmRev, vjdmFwd, vjdmRev;
  double IVjsmFwd, IVjsmRev, IVjdmFwd, IVjdmRev, SslpFwd, SslpRev, DslpFwd, DslpRev;
  double SjctTempRevSatCur, DjctTempRevSatCur, SswTempRevSatCur, DswTempRevSatCur;
  double SswgTempRevSatCur, DswgTempRevSatCur;
  /* evaluation state kept between calls and read by the loads */
  double Vd, Vs, Vb, Vsp, Vdp, Vgp, Vbp, Vge, Vgm, Vdb, Vsb, Qtotal;
  double Vddp, Vssp, Vdbb, Vdbbp, Vsbb, Vsbbp, Vbpb, Vgegp, Vgegm, Vgmgp;
  double vds, vgs, vbs, vges, vgms, vdbs, vsbs, vses, vdes, qdef;
  double vbd, vdbd, vgd, vged, vgmd, vgmb, vgb, vbs_jct, vbd_jct;
  double vbd_orig, vbs_orig, vgs_orig, vds_orig, vgd_orig, vges_orig, vgms_orig, vdes_orig;
  double vses_orig, vdbs_orig, vsbs_orig, vdbd_orig, vged_orig, vgmd_orig, vbs_jct_orig;
  double vbd_jct_orig, vgmb_orig, vgb_orig, Vgegp_orig, Vgegm_orig, Vgmgp_orig;
  int origFlag, limitedFlag, newtonIterOld, mode, ChargeComputationNeeded;
  double von, thetavth, nstar, vgs_eff, vgd_eff, dvgs_eff_dvg, dvgd_eff_dvg;
  double Vgsteff, Abulk, ueff, EsatL, vdsat, Vdseff, AbovVgst2Vtm, grdsw;
  double gbs, cbs, gbd, cbd, csub, gbbs, gbgs, gbds, cdrain, gds, gm, gmbs, IdovVds;
  double gcrg, gcrgd, gcrgg, gcrgs, gcrgb, gstot, gstotd, gstotg, gstots, gstotb;
  double gdtot, gdtotd, gdtotg, gdtots, gdtotb;
  double Igidl, ggidld, ggidlg, ggidlb, Igisl, ggisls, ggislg, ggislb, ggidls, ggisld;
  double Igcs, gIgcsg, gIgcsd, gIgcsb, gIgcss, Igcd, gIgcdg, gIgcdd, gIgcdb, gIgcds;
  double Igs, gIgsg, gIgss, Igd, gIgdg, gIgdd, Igb, gIgbg, gIgbd, gIgbs, gIgbb, cd;
  double qinv, qgate, qdrn, qsrc, qbulk, qgmid, qgdo, qgso, qgb, qgmb, qchqs, qcheq;
  double cggb, cgsb, cgdb, cdgb, cdsb, cddb, cbgb, cbsb, cbdb, csgb, cssb, csdb;
  double cgbb, csbb, cdbb, cbbb, cqdb, cqsb, cqgb, cqbb, gtau, CoxWL, taunet;
  double qbs, qbd, capbs, capbd, Igate, IgateMid, Igate_Jdxp, IgateMid_Jdxp;
  double Idrain, Isource, Idbb, Idbbp, Isbb, Isbbp, Ibpb;
  double gcggb, gcgdb, gcgsb, gcgbb, gcdgb, gcsgb, gcbgb, gcddb, gcdsb, gcsdb, gcssb;
  double gcdbb, gcsbb, gcbdb, gcbsb, gcdbdb, gcsbsb, gcbbb, gcgmgmb, gcgmdb, gcgmsb;
  double gcgmbb, gcdgmb, gcsgmb, gcbgmb, gcqgb, gcqdb, gcqsb, gcqbb, gqdef;
  double ggtg, ggtd, ggts, ggtb, gtg, gtd, gts, gtb;
  double dxpart, sxpart, ddxpart_dVd, ddxpart_dVg, ddxpart_dVb, ddxpart_dVs;
  double dsxpart_dVd, dsxpart_dVg, dsxpart_dVb, dsxpart_dVs;
  double CAPcggb, CAPcgdb, CAPcgsb, CAPcgbb, CAPcdgb, CAPcsgb, CAPcbgb, CAPcddb, CAPcdsb;
  double CAPcsdb, CAPcssb, CAPcdbb, CAPcsbb, CAPcbdb, CAPcbsb, CAPcdbdb, CAPcsbsb, CAPcbbb;
  double CAPcgmgmb, CAPcgmdb, CAPcgmsb, CAPcgmbb, CAPcdgmb, CAPcsgmb, CAPcbgmb;
  double CAPcqgb, CAPcqdb, CAPcqsb, CAPcqbb;
  double qg, qd, qb, ScalingFactor;
  double dcVgsteff, dcAbulk, dcVdseff;   /* DC-section values, for reading only */
  /* setupFVectorVars / auxChargeCalculations products */
  double Gm, Gmbs, FwdSum, RevSum, ceqdrn, ceqbd, ceqbs, ceqjs, ceqjd, ceqgcrg;
  double ceqgstot, ceqgdtot, gbbdp, gbbsp, gbdpg, gbdpdp, gbdpb, gbdpsp;
  double gbspg, gbspdp, gbspb, gbspsp;
  double gIstotg, gIstotd, gIstots, gIstotb, Istoteq, gIdtotg, gIdtotd, gIdtots, gIdtotb;
  double Idtoteq, gIbtotg, gIbtotd, gIbtots, gIbtotb, Ibtoteq, gIgtotg, gIgtotd, gIgtots;
  double gIgtotb, Igtoteq;
  double ceqdrn_Jdxp, ceqbd_Jdxp, ceqbs_Jdxp, Istoteq_Jdxp, Idtoteq_Jdxp, Ibtoteq_Jdxp;
  double Igtoteq_Jdxp, ceqgcrg_Jdxp, ceqgstot_Jdxp, ceqgdtot_Jdxp, ceqjs_Jdxp, ceqjd_Jdxp;
  double Qeqqg_Jdxp, Qeqqd_Jdxp, Qeqqb_Jdxp, Qeqqgmid_Jdxp, Qeqqjs_Jdxp, Qeqqjd_Jdxp;
  double gjbd, gjbs, gdpr, gspr, geltd;
  double Qqcheq_Jdxp;
  /* LIDs and CSR slots */
  int li_Drain, li_GateExt, li_Source, li_Body, li_DrainPrime, li_GatePrime, li_GateMid;
  int li_SourcePrime, li_BodyPrime, li_DrainBody, li_SourceBody, li_Charge;
  int li_Ibs, li_Ids, li_Igs;
  int li_store_vbd, li_store_vbs, li_store_vgs, li_store_vds, li_store_vges, li_store_vgms;
  int li_store_vdes, li_store_vses, li_store_vdbs, li_store_vsbs, li_store_vdbd;
  int li_store_vged, li_store_vgmd, li_store_von;
  int li_state_qb, li_state_qg, li_state_qd, li_state_qgmid, li_state_qbs, li_state_qbd;
  int li_state_qcheq, li_state_qcdump;
  int64_t sp[B4_NPAIR];   /* CSR slot of each (row, col) pair the loads address */
} b4inst;

#define M(x)  (m->p.x)
#define MG(x) (m->g.x)

/* setup / temperature (or_bsim4_setup.c) */
int  b4_model_init(b4model* m, const or_card* c, const xa_device_options* opt);
int  b4_instance_init(struct or_group_s* g, b4inst* I, int64_t i);
void b4_update_temperature(b4inst* I, double temp_tmp, const xa_device_options* opt);
/* evaluation (or_bsim4_eval.c) */
int  b4_update_intermediate(b4inst* I, const xa_solver_state* ss, const xa_device_options* opt,
                            const xa_vectors* v);
/* load helpers (or_bsim4_load.c) */
void b4_setup_capacitors_old(b4inst* I, const xa_solver_state* ss);
void b4_setup_capacitors_new(b4inst* I);

#endif
