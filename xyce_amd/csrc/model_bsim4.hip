// model_bsim4.hip -- BSIM4 (level 54) MOSFET: parameter tables, packing of
// the setup results into per-instance SoA columns and a per-model device
// card, and the per-instance evaluation + stamp kernel.
//
// Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET_B4.C
//   updateIntermediateVars :7744-11919, setupCapacitors_oldDAE :12616-13031,
//   setupCapacitors_newDAE :12357-12606, auxChargeCalculations :12216-12347,
//   setupFVectorVars :13525-13838, registerLIDs :4996-5128,
//   registerStateLIDs :5200-5257, registerStoreLIDs :5267-5290,
//   Master::updateState :16180-16293, Master::loadDAEVectors :16303-16567,
//   Master::loadDAEMatrices :16577-17268.
// Every branch of updateIntermediateVars is evaluated.  Groups whose cards
// or instances use mtrlMod, capMod 0/1, igcMod/igbMod, rdsMod, rgateMod,
// rbodyMod, trnqsMod/acnqsMod or initial conditions take the option kernel (k_update_x: eval with X) and the option
// stamp lists; the others take the default kernels, which compile none of
// those branches.
#include <cmath>
#include <cstring>
#include "bsim4_host.h"
#include "xa_kernel.h"

namespace xa {
namespace bsim4 {

// ---- parameter tables (names as the reference registers them) ------------
static ParamDef g_mparams[] = {
#define B4M(NAME, mem, def) {#NAME, MS_##mem, (double)(def), 0},
#include "xa_bsim4_params.def"
#undef B4M
  {"TYPE", MS_TYPE, 1.0, 0},
};
static ParamDef g_iparams[] = {
#define B4I(NAME, mem, def) {#NAME, IS_##mem, (double)(def), 0},
#include "xa_bsim4_params.def"
#undef B4I
};

// ---- device-side layout --------------------------------------------------
// Model values the evaluation reads: card parameters after defaulting and
// the temperature-dependent values the model object holds.
struct DevCard {
  double DbulkJctBotGradingCoeff;
  double DbulkJctGateSideGradingCoeff;
  double DbulkJctSideGradingCoeff;
  double DjctEmissionCoeff;
  double DjctGateSidewallTempSatCurDensity;
  double DjctSidewallTempSatCurDensity;
  double DjctTempSatCurDensity;
  double DunitAreaTempJctCap;
  double DunitLengthGateSidewallTempJctCap;
  double DunitLengthSidewallTempJctCap;
  double PhiBD;
  double PhiBS;
  double PhiBSWD;
  double PhiBSWGD;
  double PhiBSWGS;
  double PhiBSWS;
  double SbulkJctBotGradingCoeff;
  double SbulkJctGateSideGradingCoeff;
  double SbulkJctSideGradingCoeff;
  double SjctEmissionCoeff;
  double SjctGateSidewallTempSatCurDensity;
  double SjctSidewallTempSatCurDensity;
  double SjctTempSatCurDensity;
  double SunitAreaTempJctCap;
  double SunitLengthGateSidewallTempJctCap;
  double SunitLengthSidewallTempJctCap;
  double ados;
  double bdos;
  double bvd;
  double bvs;
  double coxe;
  double coxp;
  double cvchargeMod;
  double dioMod;
  double dtype;
  double epsrox;
  double factor1;
  double lambda;
  double mobMod;
  double njtsdtemp;
  double njtsstemp;
  double njtsswdtemp;
  double njtsswgdtemp;
  double njtsswgstemp;
  double njtsswstemp;
  double pditsl;
  double rdsMod;
  double tempMod;
  double tnoiMod;
  double tnom;
  double toxe;
  double toxp;
  double vcrit;
  double vtl;
  double vtm;
  double vtm0;
  double vtsd;
  double vtss;
  double vtsswd;
  double vtsswgd;
  double vtsswgs;
  double vtssws;
  double xjbvd;
  double xjbvs;
  double xpart;
  double lambdaGiven;
  double vtlGiven;
};

// Per-instance double columns: size-dependent (binned) parameters, then
// instance values fixed at setup.
#define B4_PFIELDS \
  B4P(Xdep0) \
  B4P(a0) \
  B4P(a1) \
  B4P(a2) \
  B4P(abulkCVfactor) \
  B4P(acde) \
  B4P(agidl) \
  B4P(agisl) \
  B4P(ags) \
  B4P(alpha0) \
  B4P(alpha1) \
  B4P(b0) \
  B4P(b1) \
  B4P(beta0) \
  B4P(bgidl) \
  B4P(bgisl) \
  B4P(cdep0) \
  B4P(cdsc) \
  B4P(cdscb) \
  B4P(cdscd) \
  B4P(cgbo) \
  B4P(cgdl) \
  B4P(cgdo) \
  B4P(cgidl) \
  B4P(cgisl) \
  B4P(cgsl) \
  B4P(cgso) \
  B4P(cit) \
  B4P(ckappad) \
  B4P(ckappas) \
  B4P(delta) \
  B4P(dvt0) \
  B4P(dvt0w) \
  B4P(dvt1) \
  B4P(dvt1w) \
  B4P(dvt2) \
  B4P(dvt2w) \
  B4P(dvtp0) \
  B4P(dvtp1) \
  B4P(dwb) \
  B4P(dwg) \
  B4P(egidl) \
  B4P(egisl) \
  B4P(etab) \
  B4P(eu) \
  B4P(fprout) \
  B4P(k1) \
  B4P(k1ox) \
  B4P(k3) \
  B4P(k3b) \
  B4P(keta) \
  B4P(kt1) \
  B4P(kt1l) \
  B4P(kt2) \
  B4P(lambda) \
  B4P(ldeb) \
  B4P(leff) \
  B4P(leffCV) \
  B4P(litl) \
  B4P(lpe0) \
  B4P(lpeb) \
  B4P(moin) \
  B4P(mstar) \
  B4P(mstarcv) \
  B4P(nfactor) \
  B4P(ngate) \
  B4P(noff) \
  B4P(pclm) \
  B4P(pdiblb) \
  B4P(pdits) \
  B4P(pditsd) \
  B4P(phi) \
  B4P(prwb) \
  B4P(prwg) \
  B4P(pscbe1) \
  B4P(pscbe2) \
  B4P(pvag) \
  B4P(rds0) \
  B4P(rdswmin) \
  B4P(sqrtPhi) \
  B4P(tfactor) \
  B4P(theta0vb0) \
  B4P(thetaRout) \
  B4P(ua) \
  B4P(ub) \
  B4P(uc) \
  B4P(ud) \
  B4P(vbi) \
  B4P(voffcbn) \
  B4P(voffcbncv) \
  B4P(voffcv) \
  B4P(vtl) \
  B4P(w0) \
  B4P(weff) \
  B4P(weffCJ) \
  B4P(weffCV) \
  B4P(xj) \

#define B4_VFIELDS \
  B4V(Adeff) \
  B4V(Aseff) \
  B4V(DjctTempRevSatCur) \
  B4V(DslpFwd) \
  B4V(DslpRev) \
  B4V(DswTempRevSatCur) \
  B4V(DswgTempRevSatCur) \
  B4V(IVjdmFwd) \
  B4V(IVjdmRev) \
  B4V(IVjsmFwd) \
  B4V(IVjsmRev) \
  B4V(OFF) \
  B4V(Pdeff) \
  B4V(Pseff) \
  B4V(SjctTempRevSatCur) \
  B4V(SslpFwd) \
  B4V(SslpRev) \
  B4V(SswTempRevSatCur) \
  B4V(SswgTempRevSatCur) \
  B4V(XExpBVD) \
  B4V(XExpBVS) \
  B4V(drainConductance) \
  B4V(eta0) \
  B4V(k2ox) \
  B4V(nf) \
  B4V(numberParallel) \
  B4V(sourceConductance) \
  B4V(temp) \
  B4V(u0temp) \
  B4V(vbsc) \
  B4V(vfb) \
  B4V(vfbzb) \
  B4V(vjdmFwd) \
  B4V(vjdmRev) \
  B4V(vjsmFwd) \
  B4V(vjsmRev) \
  B4V(vsattemp) \
  B4V(vtfbphi1) \
  B4V(vtfbphi2) \
  B4V(vth0) \


enum DField {
#define B4P(x) P_##x,
#define B4V(x) V_##x,
  B4_PFIELDS
  B4_VFIELDS
#undef B4P
#undef B4V
  ND
};
// Columns only the option kernel reads (gate tunnelling, bias-dependent
// source/drain resistance, capMod 0, the gate and body resistance networks,
// initial conditions, the mtrlMod material values): size-dependent values
// (B4X), instance values (B4Y) and model values (B4Z), appended after the ND
// columns for groups that use an option and always read from HBM.
#define B4_XFIELDS \
  B4X(nigc) B4X(aigc) B4X(bigc) B4X(cigc) B4X(aigs) B4X(bigs) B4X(cigs) B4X(aigd) \
  B4X(bigd) B4X(cigd) B4X(pigcd) B4X(vfbsd) B4X(vfbsdoff) B4X(Aechvb) B4X(Bechvb) \
  B4X(AechvbEdgeS) B4X(AechvbEdgeD) B4X(BechvbEdge) B4X(ToxRatio) B4X(nigbacc) \
  B4X(aigbacc) B4X(bigbacc) B4X(cigbacc) B4X(nigbinv) B4X(eigbinv) B4X(aigbinv) \
  B4X(bigbinv) B4X(cigbinv) B4X(rs0) B4X(rswmin) B4X(rd0) B4X(rdwmin) B4X(xrcrg1) \
  B4X(xrcrg2) B4X(vfbcv) \
  B4Y(grgeltd) B4Y(grbdb) B4Y(grbsb) B4Y(grbpb) B4Y(grbps) B4Y(grbpd) B4Y(icVDS) \
  B4Y(icVGS) B4Y(icVBS) \
  B4Z(eot) B4Z(epsrsub) B4Z(epsrgate) B4Z(phig) B4Z(easub) B4Z(Eg0)
enum XField {
  X_BEFORE_ = ND - 1,
#define B4X(x) X_##x,
#define B4Y(x) X_##x,
#define B4Z(x) X_##x,
  B4_XFIELDS
#undef B4X
#undef B4Y
#undef B4Z
  NDX
};
// A model's entry in the card table: the DevCard, then one value per
// per-instance column.  Columns whose value is the same for every instance
// of every card (parameters with no geometry or instance dependence, which
// the reference also shares through its size-dependent records and model)
// are read from there -- from the block's LDS copy of the cards -- instead
// of from the instance's HBM column (KArgs::ucol marks them).
constexpr int kCardBase = (int)(sizeof(DevCard) / sizeof(double));
constexpr int kCardStride = kCardBase + ND;
// The kernels take the card-uniform columns as a compile-time set, so that a
// parameter read is a plain LDS or HBM load with no per-read branch (one in
// every read measurably worsened register allocation).  CM (card mode):
//   0: every column from HBM;
//   1: the size-dependent columns except the geometry-derived ones (uniform
//      for binless cards whatever the instance parameters: the Monte Carlo
//      rings with per-instance DELVTO, the netlists);
//   2: set 1 plus the instance-value columns except those of the drain/source
//      perimeters and vfbzb (also uniform when instances give no AS/AD/NRD/
//      NRS/DELVTO/TEMP/M and share a temperature: the microbenchmark).
// The launch takes the largest set that the group's runtime mask (ucol)
// covers, so the bits read are the same in every mode.
constexpr int kGeomP[] = {P_abulkCVfactor, P_cgdo, P_cgso, P_leff, P_leffCV, P_rds0,
                          P_theta0vb0, P_thetaRout, P_weff, P_weffCJ, P_weffCV};
constexpr int kGeomV[] = {V_Pdeff, V_Pseff, V_vfbzb};
// the size-dependent columns come first in DField, then the instance values
enum {
#define B4P(x) C_##x,
#define B4V(x)
  B4_PFIELDS
#undef B4P
#undef B4V
  kNumP
};
__host__ __device__ constexpr bool in_list(const int* l, int m, int f) {
  for (int k = 0; k < m; ++k)
    if (l[k] == f) return true;
  return false;
}
// column f is read from the card in card mode cm
__host__ __device__ constexpr bool card_col(int cm, int f) {
  return cm == 0 ? false
                 : f < kNumP ? !in_list(kGeomP, (int)(sizeof(kGeomP) / sizeof(int)), f)
                             : cm == 2 && !in_list(kGeomV, (int)(sizeof(kGeomV) / sizeof(int)), f);
}
static_assert(P_a0 < kNumP && V_vth0 >= kNumP, "size-dependent columns precede instance values");

enum IField { F_MODEL, F_DEXIST, F_SEXIST, NIF, F_XF = NIF };
// Option bits of an option group's instance (ifield F_XF): the instance's
// rgateMod, rbodyMod, trnqsMod and acnqsMod, its model's igcMod, igbMod,
// capMod, rdsMod and rbodyMod (the reference tests the model's value for the
// body currents), PIGCD given, and the initial conditions given.
enum XBit {
  XB_RGATE = 0,      // 2 bits
  XB_RBODY = 2,      // 2 bits
  XB_TRNQS = 4, XB_ACNQS = 5,
  XB_IGC = 6,        // 2 bits
  XB_IGB = 8,
  XB_CAP = 9,        // 2 bits
  XB_RDS = 11, XB_RBODYM = 12, XB_PIGCD = 13, XB_ICVBS = 14, XB_ICVDS = 15, XB_ICVGS = 16,
  XB_MTRL = 17
};
// Values the reference's Instance keeps from one evaluation to the next and
// reads before writing: vgb (as vgb_orig), and the junction charges and
// capacitances, which are only recomputed when charges are needed (transient,
// tranop, acop) and are otherwise the last computed ones.
enum State { S_VGB, S_QBS, S_QBD, S_CAPBS, S_CAPBD, NS };
static const char* const kStateNames[NS] = {"vgb", "qbs", "qbd", "capbs", "capbd"};
// Local nodes: the four terminals (N_G is the external gate), the drain and
// source primes (the terminals themselves when no internal node exists), the
// gate prime and gate mid nodes (rgateMod > 0, rgateMod = 3; else the
// external gate), the body prime, source-body and drain-body nodes (rbodyMod;
// else the body), the NQS charge node (trnqsMod) and the branch currents of
// the initial conditions IC3 (VBS), IC1 (VDS), IC2 (VGS) (absent: -1).  The
// default kernel reads the first six only.
enum Node { N_D, N_G, N_S, N_B, N_DP, N_SP, N_GP, N_GM, N_BP, N_SB, N_DB, N_Q, N_IBS, N_IDS,
            N_IGS, NNODE };
enum Sto {
  T_vbd, T_vbs, T_vgs, T_vds, T_vges, T_vgms, T_vdes, T_vses, T_vdbs, T_vsbs,
  T_vdbd, T_vged, T_vgmd, T_von, NSTO
};
// State slots: qb, qg, qd, then as the options register them qgmid
// (rgateMod 3), qbs and qbd (rbodyMod), qcheq and qcdump (trnqsMod), so an
// option instance's slot of each follows from its flags.
enum Sta { A_qb, A_qg, A_qd, NSTA_BASE, NSTA = 8 };
// Entries per instance of the option groups' stamp lists (F, Q, F limiter,
// Q limiter, dF/dx, dQ/dx)
constexpr int kXF = 27, kXQ = 8, kXFLim = 13, kXQLim = 8, kXDF = 99, kXDQ = 33;

// ---- host setup ------------------------------------------------------------
static DevCard make_card(const Model& m) {
  DevCard c{};
    c.DbulkJctBotGradingCoeff = m.p.DbulkJctBotGradingCoeff;
    c.DbulkJctGateSideGradingCoeff = m.p.DbulkJctGateSideGradingCoeff;
    c.DbulkJctSideGradingCoeff = m.p.DbulkJctSideGradingCoeff;
    c.DjctEmissionCoeff = m.p.DjctEmissionCoeff;
    c.DjctGateSidewallTempSatCurDensity = m.DjctGateSidewallTempSatCurDensity;
    c.DjctSidewallTempSatCurDensity = m.DjctSidewallTempSatCurDensity;
    c.DjctTempSatCurDensity = m.DjctTempSatCurDensity;
    c.DunitAreaTempJctCap = m.DunitAreaTempJctCap;
    c.DunitLengthGateSidewallTempJctCap = m.DunitLengthGateSidewallTempJctCap;
    c.DunitLengthSidewallTempJctCap = m.DunitLengthSidewallTempJctCap;
    c.PhiBD = m.PhiBD;
    c.PhiBS = m.PhiBS;
    c.PhiBSWD = m.PhiBSWD;
    c.PhiBSWGD = m.PhiBSWGD;
    c.PhiBSWGS = m.PhiBSWGS;
    c.PhiBSWS = m.PhiBSWS;
    c.SbulkJctBotGradingCoeff = m.p.SbulkJctBotGradingCoeff;
    c.SbulkJctGateSideGradingCoeff = m.p.SbulkJctGateSideGradingCoeff;
    c.SbulkJctSideGradingCoeff = m.p.SbulkJctSideGradingCoeff;
    c.SjctEmissionCoeff = m.p.SjctEmissionCoeff;
    c.SjctGateSidewallTempSatCurDensity = m.SjctGateSidewallTempSatCurDensity;
    c.SjctSidewallTempSatCurDensity = m.SjctSidewallTempSatCurDensity;
    c.SjctTempSatCurDensity = m.SjctTempSatCurDensity;
    c.SunitAreaTempJctCap = m.SunitAreaTempJctCap;
    c.SunitLengthGateSidewallTempJctCap = m.SunitLengthGateSidewallTempJctCap;
    c.SunitLengthSidewallTempJctCap = m.SunitLengthSidewallTempJctCap;
    c.ados = m.p.ados;
    c.bdos = m.p.bdos;
    c.bvd = m.p.bvd;
    c.bvs = m.p.bvs;
    c.coxe = m.coxe;
    c.coxp = m.coxp;
    c.cvchargeMod = m.p.cvchargeMod;
    c.dioMod = m.p.dioMod;
    c.dtype = m.dtype;
    c.epsrox = m.p.epsrox;
    c.factor1 = m.factor1;
    c.lambda = m.p.lambda;
    c.mobMod = m.p.mobMod;
    c.njtsdtemp = m.njtsdtemp;
    c.njtsstemp = m.njtsstemp;
    c.njtsswdtemp = m.njtsswdtemp;
    c.njtsswgdtemp = m.njtsswgdtemp;
    c.njtsswgstemp = m.njtsswgstemp;
    c.njtsswstemp = m.njtsswstemp;
    c.pditsl = m.p.pditsl;
    c.rdsMod = m.p.rdsMod;
    c.tempMod = m.p.tempMod;
    c.tnoiMod = m.p.tnoiMod;
    c.tnom = m.p.tnom;
    c.toxe = m.p.toxe;
    c.toxp = m.p.toxp;
    c.vcrit = m.vcrit;
    c.vtl = m.p.vtl;
    c.vtm = m.vtm;
    c.vtm0 = m.vtm0;
    c.vtsd = m.p.vtsd;
    c.vtss = m.p.vtss;
    c.vtsswd = m.p.vtsswd;
    c.vtsswgd = m.p.vtsswgd;
    c.vtsswgs = m.p.vtsswgs;
    c.vtssws = m.p.vtssws;
    c.xjbvd = m.p.xjbvd;
    c.xjbvs = m.p.xjbvs;
    c.xpart = m.p.xpart;
    c.lambdaGiven = m.g.lambda ? 1.0 : 0.0;
    c.vtlGiven = m.g.vtl ? 1.0 : 0.0;
  return c;
}

// model values of the mtrlMod columns (the model's Eg0 is its temperature value)
static double mtrl_value(const Model& m, const char* x) {
  if (!std::strcmp(x, "Eg0")) return m.Eg0;
  if (!std::strcmp(x, "eot")) return m.p.eot;
  if (!std::strcmp(x, "epsrsub")) return m.p.epsrsub;
  if (!std::strcmp(x, "epsrgate")) return m.p.epsrgate;
  if (!std::strcmp(x, "phig")) return m.p.phig;
  return m.p.easub;
}

static int setup(Group& g) {
  const int64_t n = g.n;
  std::vector<Model> models(g.models.size());
  for (size_t k = 0; k < g.models.size(); ++k) {
    if (g.models[k].level != 54 && g.models[k].level != 14)
      return fail(XA_ERR_UNSUPPORTED, "bsim4: level must be 14 or 54");
    const int rc = model_init(models[k], g.models[k], g.opt);
    if (rc != XA_OK) return rc;
  }
  // An option group (any card or instance with an option outside the
  // default kernel's branches) takes the option kernel and its longer stamp
  // lists, the extra columns and the option bits.
  bool xopt = false;
  for (const Model& m : models)
    xopt = xopt || m.p.mtrlMod != 0 || m.p.capMod != 2 || m.p.igcMod != 0 || m.p.igbMod != 0 || m.p.rdsMod != 0 ||
           m.p.rbodyMod != 0 || m.p.rgateMod != 0 || m.p.trnqsMod != 0 || m.p.acnqsMod != 0;
  for (int s : {IS_rgateMod, IS_rbodyMod, IS_trnqsMod, IS_acnqsMod})
    if (g.igiven[s])
      for (int64_t i = 0; i < n && !xopt; ++i) xopt = g.ival[s][i] != 0.0;
  for (int s : {IS_icVDS, IS_icVGS, IS_icVBS}) xopt = xopt || g.igiven[s];
  const int ncol = xopt ? NDX : ND;
  g.variant = xopt ? 1 : 0;
  g.split_only = xopt;
  if (xopt) {
    const int E[NK] = {kXF, kXQ, kXFLim, kXQLim, kXDF, kXDQ, 0, 0};
    for (int k = 0; k < NK; ++k) g.ent[k] = E[k];
  }
  g.n_dfield = ncol;
  g.n_ifield = xopt ? NIF + 1 : NIF;
  g.dfield.assign((size_t)ncol * n, 0.0);
  g.ifield.assign((size_t)g.n_ifield * n, 0);
  g.n_state = NS;
  g.state_init.assign((size_t)NS * n, 0.0);
  auto D = [&](int f, int64_t i) -> double& { return g.dfield[f * n + i]; };
  auto I = [&](int f, int64_t i) -> int32_t& { return g.ifield[f * n + i]; };
  std::unique_ptr<Inst> rec(new Inst());
  for (int64_t i = 0; i < n; ++i) {
    Inst& in = *rec;
    in = Inst();
    in.model = &models[g.inst_model[i]];
    const double l = g.ival[IS_l][i], w = g.ival[IS_w][i], nf = g.ival[IS_nf][i];
    if (std::isnan(l) || std::isnan(w) || std::isnan(nf))
      return fail(XA_ERR_ARG, "bsim4: L, W and NF must be numbers");
    const int rc = instance_init(g, in, i);
    if (rc != XA_OK) return rc;
    const Size& p = *in.paramPtr;
#define B4P(x) D(P_##x, i) = p.x;
#define B4V(x) D(V_##x, i) = in.x;
    B4_PFIELDS
    B4_VFIELDS
#undef B4P
#undef B4V
    I(F_MODEL, i) = g.inst_model[i];
    I(F_DEXIST, i) = in.drainMOSFET_B4Exists;
    I(F_SEXIST, i) = in.sourceMOSFET_B4Exists;
    if (xopt) {
#define B4X(x) D(X_##x, i) = p.x;
#define B4Y(x) D(X_##x, i) = in.x;
#define B4Z(x) D(X_##x, i) = mtrl_value(*in.model, #x);
      B4_XFIELDS
#undef B4X
#undef B4Y
#undef B4Z
      const Model& m = *in.model;
      I(F_XF, i) = (in.irgateMod << XB_RGATE) | (in.irbodyMod << XB_RBODY) |
                   ((in.itrnqsMod ? 1 : 0) << XB_TRNQS) | ((in.iacnqsMod ? 1 : 0) << XB_ACNQS) |
                   ((int)m.p.igcMod << XB_IGC) | ((int)m.p.igbMod << XB_IGB) |
                   ((int)m.p.capMod << XB_CAP) | ((m.p.rdsMod != 0 ? 1 : 0) << XB_RDS) |
                   ((m.p.rbodyMod != 0 ? 1 : 0) << XB_RBODYM) | ((m.g.pigcd ? 1 : 0) << XB_PIGCD) |
                   ((in.icVBS_given ? 1 : 0) << XB_ICVBS) | ((in.icVDS_given ? 1 : 0) << XB_ICVDS) |
                   ((in.icVGS_given ? 1 : 0) << XB_ICVGS) |
                   ((m.p.mtrlMod != 0 ? 1 : 0) << XB_MTRL);
    }
  }
  // cards after every instance: the model's temperature values are those
  // the last instance's updateTemperature wrote
  g.mcard_stride = kCardStride;
  g.mcard.assign(g.models.size() * g.mcard_stride, 0.0);
  for (size_t k = 0; k < models.size(); ++k) {
    const DevCard c = make_card(models[k]);
    std::memcpy(&g.mcard[k * g.mcard_stride], &c, sizeof(c));
  }
  // card-uniform columns: every instance of each card holds the same bits
  std::vector<int64_t> first(models.size(), -1);
  for (int64_t i = 0; i < n; ++i)
    if (first[g.inst_model[i]] < 0) first[g.inst_model[i]] = i;
  g.ucol[0] = g.ucol[1] = g.ucol[2] = 0;
  for (int f = 0; f < ND; ++f) {
    bool uniform = true;
    const double* col = &g.dfield[(size_t)f * n];
    for (int64_t i = 0; i < n && uniform; ++i)
      uniform = std::memcmp(&col[i], &col[first[g.inst_model[i]]], sizeof(double)) == 0;
    if (!uniform) continue;
    g.ucol[f >> 6] |= 1ull << (f & 63);
    for (size_t k = 0; k < models.size(); ++k)
      if (first[k] >= 0) g.mcard[k * g.mcard_stride + kCardBase + f] = col[first[k]];
  }
  return XA_OK;
}

// option bits of instance i of an option group (0 in a default group)
static int xflags(const Group& g, int64_t i) {
  return g.variant ? g.ifield[(size_t)F_XF * g.n + i] : 0;
}

// registerLIDs (N_DEV_MOSFET_B4.C:4996-5128): external nodes, then the
// internal ones in the reference's order.
static void resolve_nodes(const Group& g, int64_t i, const int32_t* reg, int32_t* full) {
  const int64_t n = g.n;
  const int xf = xflags(g, i);
  const int rgate = (xf >> XB_RGATE) & 3, rbody = (xf >> XB_RBODY) & 3;
  full[N_D] = reg[0];
  full[N_G] = reg[1];
  full[N_S] = reg[2];
  full[N_B] = reg[3];
  int loc = 4;
  full[N_DP] = g.ifield[F_DEXIST * n + i] ? reg[loc++] : reg[0];
  full[N_SP] = g.ifield[F_SEXIST * n + i] ? reg[loc++] : reg[2];
  full[N_GP] = rgate > 0 ? reg[loc++] : reg[1];
  full[N_GM] = rgate == 3 ? reg[loc++] : reg[1];
  if (rbody) {
    full[N_BP] = reg[loc++];
    full[N_SB] = reg[loc++];
    full[N_DB] = reg[loc++];
  } else {
    full[N_BP] = full[N_SB] = full[N_DB] = reg[3];
  }
  full[N_Q] = ((xf >> XB_TRNQS) & 1) ? reg[loc++] : -1;
  full[N_IBS] = ((xf >> XB_ICVBS) & 1) ? reg[loc++] : -1;
  full[N_IDS] = ((xf >> XB_ICVDS) & 1) ? reg[loc++] : -1;
  full[N_IGS] = ((xf >> XB_ICVGS) & 1) ? reg[loc++] : -1;
}

// Stamp tables in the order of the reference's loads.  A row of -1 marks an
// entry the instance does not load (no internal drain/source node).
constexpr int16_t N_NONE = -1;
static const Stamp kStF[] = {{N_DP, 0}, {N_G, 0}, {N_B, 0}, {N_SP, 0}, {N_D, 0}, {N_DP, 0}, {N_S, 0}, {N_SP, 0}};
static const Stamp kStQ[] = {{N_DP, 0}, {N_G, 0}, {N_B, 0}, {N_SP, 0}};
static const Stamp kStLim[] = {{N_DP, 0}, {N_G, 0}, {N_B, 0}, {N_SP, 0}};
static const Stamp kStDF[] = {{N_G, N_G}, {N_G, N_DP}, {N_G, N_SP}, {N_G, N_B}, {N_DP, N_DP}, {N_DP, N_D}, {N_DP, N_G}, {N_DP, N_SP}, {N_DP, N_B}, {N_D, N_DP}, {N_D, N_D}, {N_SP, N_DP}, {N_SP, N_G}, {N_SP, N_SP}, {N_SP, N_S}, {N_SP, N_B}, {N_S, N_SP}, {N_S, N_S}, {N_B, N_DP}, {N_B, N_G}, {N_B, N_SP}, {N_B, N_B}, {N_DP, N_DP}, {N_DP, N_G}, {N_DP, N_SP}, {N_DP, N_B}, {N_B, N_DP}, {N_B, N_G}, {N_B, N_SP}, {N_B, N_B}, {N_SP, N_DP}, {N_SP, N_G}, {N_SP, N_SP}, {N_SP, N_B}, {N_B, N_DP}, {N_B, N_G}, {N_B, N_SP}, {N_B, N_B}};
static const Stamp kStDQ[] = {{N_G, N_G}, {N_G, N_DP}, {N_G, N_SP}, {N_G, N_B}, {N_DP, N_DP}, {N_DP, N_G}, {N_DP, N_SP}, {N_DP, N_B}, {N_SP, N_DP}, {N_SP, N_G}, {N_SP, N_SP}, {N_SP, N_B}, {N_B, N_DP}, {N_B, N_G}, {N_B, N_SP}, {N_B, N_B}};

// Option groups: one stamp list per kind holding every entry of every option
// branch of the reference's loads in its order; an instance disables the
// entries of branches its options do not take (row N_NONE), which keeps the
// relative order of the entries it loads.
constexpr int16_t GE = N_G, GP = N_GP, GM = N_GM, DP = N_DP, SP = N_SP, BP = N_BP, DB = N_DB,
                  SB = N_SB, QN = N_Q, ND_ = N_D, NS_ = N_S, NB_ = N_B;
static const Stamp kXStF[] = {
    {DP, 0}, {GP, 0},                                             // 0-1
    {GE, 0}, {GM, 0}, {GP, 0},                                    // 2-4 gate resistance
    {DB, 0}, {BP, 0}, {NB_, 0}, {SB, 0}, {SP, 0},                 // 5-9 body
    {ND_, 0}, {NS_, 0}, {DP, 0}, {SP, 0},                         // 10-13 rdsMod
    {ND_, 0}, {DP, 0}, {NS_, 0}, {SP, 0},                         // 14-17 contact resistors
    {NB_, 0}, {NS_, 0}, {N_IBS, 0}, {ND_, 0}, {NS_, 0}, {N_IDS, 0},
    {GE, 0}, {NS_, 0}, {N_IGS, 0}};                               // 18-26 initial conditions
static const Stamp kXStFLim[] = {{DP, 0}, {GP, 0}, {GE, 0}, {GM, 0}, {GP, 0}, {DB, 0}, {BP, 0},
                                 {SB, 0}, {SP, 0}, {ND_, 0}, {NS_, 0}, {DP, 0}, {SP, 0}};
static const Stamp kXStQ[] = {{DP, 0}, {GP, 0}, {GM, 0}, {DB, 0}, {BP, 0}, {SB, 0}, {SP, 0}, {QN, 0}};
static const Stamp kXStDF[] = {
    {GE, GE}, {GE, GP}, {GE, GM}, {GE, DP}, {GE, SP}, {GE, BP},            // 0-5
    {GM, GE}, {GM, GM}, {GM, DP}, {GM, GP}, {GM, SP}, {GM, BP},            // 6-11
    {GP, GE}, {GP, GM}, {GP, GP}, {GP, DP}, {GP, SP}, {GP, BP},            // 12-17
    {ND_, GP}, {ND_, SP}, {ND_, BP}, {NS_, DP}, {NS_, GP}, {NS_, BP},      // 18-23 rdsMod
    {DP, DP}, {DP, ND_}, {DP, GP}, {DP, SP}, {DP, BP}, {ND_, DP}, {ND_, ND_},
    {SP, DP}, {SP, GP}, {SP, SP}, {SP, NS_}, {SP, BP}, {NS_, SP}, {NS_, NS_},
    {BP, DP}, {BP, GP}, {BP, SP}, {BP, BP},                                // 24-41
    {DP, DP}, {DP, GP}, {DP, SP}, {DP, BP}, {BP, DP}, {BP, GP}, {BP, SP}, {BP, BP},  // 42-49 GIDL
    {SP, DP}, {SP, GP}, {SP, SP}, {SP, BP}, {BP, DP}, {BP, GP}, {BP, SP}, {BP, BP},  // 50-57 GISL
    {DP, DB}, {SP, SB}, {DB, DP}, {DB, DB}, {DB, BP}, {DB, NB_}, {BP, DB}, {BP, NB_},
    {BP, SB}, {BP, BP}, {SB, SP}, {SB, BP}, {SB, NB_}, {SB, SB}, {NB_, DB}, {NB_, BP},
    {NB_, SB}, {NB_, NB_},                                                 // 58-75 body
    {QN, QN}, {QN, GP}, {QN, DP}, {QN, SP}, {QN, BP}, {DP, QN}, {SP, QN}, {GP, QN},  // 76-83 NQS
    {NB_, N_IBS}, {NS_, N_IBS}, {N_IBS, NB_}, {N_IBS, NS_}, {N_IBS, N_IBS},
    {ND_, N_IDS}, {NS_, N_IDS}, {N_IDS, ND_}, {N_IDS, NS_}, {N_IDS, N_IDS},
    {GE, N_IGS}, {NS_, N_IGS}, {N_IGS, GE}, {N_IGS, NS_}, {N_IGS, N_IGS}};   // 84-98 IC
static const Stamp kXStDQ[] = {
    {GM, GM}, {GM, DP}, {GM, SP}, {GM, BP}, {DP, GM}, {SP, GM}, {BP, GM},  // 0-6 rgateMod 3
    {GP, GP}, {GP, DP}, {GP, SP}, {GP, BP}, {DP, DP}, {DP, GP}, {DP, SP}, {DP, BP},
    {SP, DP}, {SP, GP}, {SP, SP}, {SP, BP}, {BP, DP}, {BP, GP}, {BP, SP}, {BP, BP},  // 7-22
    {DP, DB}, {SP, SB}, {DB, DP}, {DB, DB}, {SB, SP}, {SB, SB},            // 23-28 body
    {QN, GP}, {QN, DP}, {QN, SP}, {QN, BP}};                               // 29-32 NQS
static_assert(sizeof(kXStF) / sizeof(Stamp) == kXF && sizeof(kXStFLim) / sizeof(Stamp) == kXFLim &&
              sizeof(kXStQ) / sizeof(Stamp) == kXQ && sizeof(kXStDF) / sizeof(Stamp) == kXDF &&
              sizeof(kXStDQ) / sizeof(Stamp) == kXDQ, "option stamp tables");

// the entries of the option stamp lists an instance with option bits xf loads
static bool xstamp_on(int kind, int e, int xf) {
  const int rgate = (xf >> XB_RGATE) & 3;
  const bool rbody = ((xf >> XB_RBODY) & 3) != 0, nqs = (xf >> XB_TRNQS) & 1;
  const bool rds = (xf >> XB_RDS) & 1;
  switch (kind) {
    case K_F:
      if (e == 2 || e == 4) return rgate >= 1;
      if (e == 3) return rgate == 3;
      if (e == 5 || e == 7 || e == 8) return rbody;
      if (e >= 10 && e <= 13) return rds;
      if (e >= 18 && e <= 20) return (xf >> XB_ICVBS) & 1;
      if (e >= 21 && e <= 23) return (xf >> XB_ICVDS) & 1;
      if (e >= 24) return (xf >> XB_ICVGS) & 1;
      return true;   // 14-17 by the contact resistors' nodes, below
    case K_FLIM:
      if (e == 2 || e == 4) return rgate >= 1;
      if (e == 3) return rgate == 3;
      if (e == 5 || e == 7) return rbody;
      if (e >= 9) return rds;
      return true;
    case K_Q: case K_QLIM:
      if (e == 2) return rgate == 3;
      if (e == 3 || e == 5) return rbody;
      if (e == 7) return nqs;
      return true;
    case K_DFDX:
      if (e == 0) return rgate >= 1;
      if (e == 1) return rgate == 1 || rgate == 2;
      if (e == 2) return rgate == 3;
      if (e >= 3 && e <= 5) return rgate == 2;
      if (e >= 6 && e <= 11) return rgate == 3;
      if (e == 12) return rgate == 1 || rgate == 2;
      if (e == 13) return rgate == 3;
      if (e >= 18 && e <= 23) return rds;
      if (e >= 58 && e <= 75) return rbody;
      if (e >= 76 && e <= 83) return nqs;
      if (e >= 84 && e <= 88) return (xf >> XB_ICVBS) & 1;
      if (e >= 89 && e <= 93) return (xf >> XB_ICVDS) & 1;
      if (e >= 94) return (xf >> XB_ICVGS) & 1;
      return true;
    case K_DQDX:
      if (e <= 6) return rgate == 3;
      if (e >= 23 && e <= 28) return rbody;
      if (e >= 29) return nqs;
      return true;
    default: return false;
  }
}

static void stamps(const Group& g, int64_t i, int kind, std::vector<Stamp>& out) {
  out.clear();
  auto add = [&](const Stamp* s, int m) { out.assign(s, s + m); };
  if (g.variant) {
    const int xf = xflags(g, i);
    switch (kind) {
      case K_F: add(kXStF, kXF); break;
      case K_Q: add(kXStQ, kXQ); break;
      case K_FLIM: if (g.opt.voltageLimiterFlag) add(kXStFLim, kXFLim); break;
      case K_QLIM: if (g.opt.voltageLimiterFlag) add(kXStQ, kXQLim); break;
      case K_DFDX: add(kXStDF, kXDF); break;
      case K_DQDX: add(kXStDQ, kXDQ); break;
      default: break;
    }
    for (size_t e = 0; e < out.size(); ++e)
      if (!xstamp_on(kind, (int)e, xf)) out[e].row = N_NONE;
    if (kind == K_F) {
      const int64_t n = g.n;
      if (!g.ifield[F_DEXIST * n + i]) out[14].row = out[15].row = N_NONE;
      if (!g.ifield[F_SEXIST * n + i]) out[16].row = out[17].row = N_NONE;
    }
    return;
  }
  switch (kind) {
    case K_F: {
      add(kStF, 8);
      const int64_t n = g.n;
      if (!g.ifield[F_DEXIST * n + i]) out[4].row = out[5].row = N_NONE;
      if (!g.ifield[F_SEXIST * n + i]) out[6].row = out[7].row = N_NONE;
      break;
    }
    case K_Q: add(kStQ, 4); break;
    case K_FLIM: case K_QLIM: if (g.opt.voltageLimiterFlag) add(kStLim, 4); break;
    case K_DFDX: add(kStDF, 38); break;
    case K_DQDX: add(kStDQ, 16); break;
    default: break;   // the BSIM4 Master loads store no lead currents
  }
}

// dQdx is not loaded on the first iteration of a transient step
// (Master::loadDAEMatrices, N_DEV_MOSFET_B4.C:17230-17236).
static bool skip_kind(const Group&, const xa_solver_state& ss, int kind) {
  return kind == K_DQDX && !ss.dcopFlag && ss.initTranFlag && ss.newtonIter == 0;
}

// ---- device evaluation ---------------------------------------------------
// DEXP of the reference: exp with linear extrapolation above the threshold.
XA_HD void dexp3(double A, double& B, double& dB) {
  if (A > kExpThreshold) { B = kMaxExp * (1.0 + A - kExpThreshold); dB = kMaxExp; }
  else if (A < -kExpThreshold) { B = kMinExp; dB = 0; }
  else { B = exp(A); dB = B; }
}

// polyDepletion (N_DEV_MOSFET_B4.C:14252-14282)
XA_HD void polyDepletion(double phi, double ngate, double epsgate, double coxe, double Vgs_arg,
                         double& Vgs_eff, double& dVgs_eff_dVg) {
  if ((ngate > 1.0e18) && (ngate < 1.0e25) && (Vgs_arg > phi) && (epsgate != 0)) {
    const double T1 = 1.0e6 * kQ * epsgate * ngate / (coxe * coxe);
    const double T8 = Vgs_arg - phi;
    const double T4 = sqrt(1.0 + 2.0 * T8 / T1);
    const double T2 = 2.0 * T8 / (T4 + 1.0);
    const double T3 = 0.5 * T2 * T2 / T1;
    const double T7 = 1.12 - T3 - 0.05;
    const double T6 = sqrt(T7 * T7 + 0.224);
    const double T5 = 1.12 - 0.5 * (T7 + T6);
    Vgs_eff = Vgs_arg - T5;
    dVgs_eff_dVg = 1.0 - (0.5 - 0.5 / T4) * (1.0 + T7 / T6);
  } else {
    Vgs_eff = Vgs_arg;
    dVgs_eff_dVg = 1.0;
  }
}

// store-vector value, 0 for an unregistered slot
XA_HD double stv(const double* s, int32_t lid) { return lid >= 0 ? s[lid] : 0.0; }

// per-instance parameter columns, read through `dfb` (see reload_params)
// (card-uniform columns of card mode CM from the card's copy, see card_col)
#define PP(x) (card_col(CM, P_##x) ? Ccol[P_##x] : dfb[(int64_t)(P_##x) * a.n + i])
#define IV(x) (card_col(CM, V_##x) ? Ccol[V_##x] : dfld(a, V_##x, i))

// A section boundary for the per-instance parameter columns: the compiler
// takes the column base as changed, so a parameter used in a later section is
// read again there (from L1/L2, or from the LDS card for a card-uniform
// column) instead of being held in registers across sections.  (The base is
// a vector operand: the compiler may fold a column read and a card read into
// one per-lane address.)  The eval's register peak (the DC model around Vasat) holds
// fewer long-lived parameters, and spills less.
__device__ __forceinline__ void reload_params(const double*& dfb) {
  asm volatile("" : "+v"(dfb));
}

// LDSCARD: the block's model cards are staged in LDS by the kernel (s_card,
// in the order of the block's card table; card_slot gives the instance's);
// otherwise they are read from global memory.
// X: the option kernel (option groups): every option branch of the
// reference is evaluated and the stamps follow the option stamp lists.  The
// default kernel (X false) compiles none of them.
#define XP(x) (dfb[(int64_t)(X_##x) * a.n + i])
template <int MODE, bool LDSCARD, int CM, bool RELOAD = true, bool X = false>
__device__ __forceinline__ void eval(const KArgs& a, int64_t i, const double* s_card) {
  if (i >= a.n) return;
  // the fused kernel stages every card of the group (slot = model index),
  // the stored and pattern kernels their block's card table
  const DevCard& C =
      LDSCARD ? *reinterpret_cast<const DevCard*>(
                    s_card + (MODE == 1 ? ifld(a, F_MODEL, i) : (int)a.card_slot[i]) * a.mstride)
              : *reinterpret_cast<const DevCard*>(a.mcard + (int64_t)ifld(a, F_MODEL, i) * a.mstride);
  const double* Ccol = reinterpret_cast<const double*>(&C) + kCardBase;
  const double* dfb = a.dfield;
  const xa_solver_state& ss = a.ss;
  const xa_device_options& opt = a.opt;
  const double gmin = opt.gmin;
  const double nf = IV(nf);
  const int xf = X ? ifld(a, F_XF, i) : 0;
  const int irgateMod = (xf >> XB_RGATE) & 3, irbodyMod = (xf >> XB_RBODY) & 3;
  const int itrnqsMod = (xf >> XB_TRNQS) & 1, iacnqsMod = (xf >> XB_ACNQS) & 1;
  const int igcMod = (xf >> XB_IGC) & 3, igbMod = (xf >> XB_IGB) & 1;
  const int capMod = X ? (xf >> XB_CAP) & 3 : 2, rdsMod = (xf >> XB_RDS) & 1;
  const int mtrlMod = (xf >> XB_MTRL) & 1;
  const int32_t li_Drain = node(a, N_D, i), li_GateExt = node(a, N_G, i);
  const int32_t li_Source = node(a, N_S, i), li_Body = node(a, N_B, i);
  const int32_t li_DrainPrime = node(a, N_DP, i), li_SourcePrime = node(a, N_SP, i);
  const int32_t li_GatePrime = X ? node(a, N_GP, i) : li_GateExt;
  const int32_t li_GateMid = X ? node(a, N_GM, i) : li_GateExt;
  const int32_t li_BodyPrime = X ? node(a, N_BP, i) : li_Body;
  const int32_t li_DrainBody = X ? node(a, N_DB, i) : li_Body;
  const int32_t li_SourceBody = X ? node(a, N_SB, i) : li_Body;
  const int32_t li_Charge = X ? node(a, N_Q, i) : -1;
  const int32_t li_store_vbd = stolid(a, T_vbd, i), li_store_vbs = stolid(a, T_vbs, i);
  const int32_t li_store_vgs = stolid(a, T_vgs, i), li_store_vds = stolid(a, T_vds, i);
  const int32_t li_store_vges = stolid(a, T_vges, i), li_store_vgms = stolid(a, T_vgms, i);
  const int32_t li_store_vdes = stolid(a, T_vdes, i), li_store_vses = stolid(a, T_vses, i);
  const int32_t li_store_vdbs = stolid(a, T_vdbs, i), li_store_vsbs = stolid(a, T_vsbs, i);
  const int32_t li_store_vdbd = stolid(a, T_vdbd, i), li_store_vged = stolid(a, T_vged, i);
  const int32_t li_store_vgmd = stolid(a, T_vgmd, i), li_store_von = stolid(a, T_von, i);

  // evaluation state of the instance
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
  double cgbb, csbb, cdbb, cbbb, cqdb, cqsb, cqgb, cqbb, CoxWL, taunet;
  double gtau = 0.0;   // no NQS: no NQS time constant
  double Qqcheq_Jdxp = 0.0;
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
  double qg, qd, qb;
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
  double cgdo, cgso;
  // vgb keeps the previous evaluation's value until it is recomputed
  vgb = st(a, S_VGB, i);
  if (!(ss.tranopFlag || ss.acopFlag || ss.transientFlag)) {
    qbs = st(a, S_QBS, i);
    qbd = st(a, S_QBD, i);
    capbs = st(a, S_CAPBS, i);
    capbd = st(a, S_CAPBD, i);
  }

  // ---- Instance::updateIntermediateVars ----

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
  double vdbs_old, vsbs_old, vdbd_old, vged_old, vgmd_old;
  int Check = 0;

  dT1_dVb = dT2_dVb = dT3_dVb = dT4_dVb = dT5_dVb = dT6_dVb = 0.0;
  dT0_dVb = 0.0;
  devbs_dvb = devbd_dvb = 0.0;
  ChargeComputationNeeded = (ss.tranopFlag || ss.acopFlag || ss.transientFlag) ? 1 : 0;
  limitedFlag = 0;

  if (RELOAD) reload_params(dfb);
  /* node voltages (:7919-7974) */
  Vd = solv(a.v.nextSol, li_Drain);
  Vs = solv(a.v.nextSol, li_Source);
  Vb = solv(a.v.nextSol, li_Body);
  Vsp = solv(a.v.nextSol, li_SourcePrime);
  Vdp = solv(a.v.nextSol, li_DrainPrime);
  Vgp = solv(a.v.nextSol, li_GatePrime);
  Vbp = solv(a.v.nextSol, li_BodyPrime);
  Vge = solv(a.v.nextSol, li_GateExt);
  Vgm = (li_GateMid >= 0) ? solv(a.v.nextSol, li_GateMid) : 0.0;
  Vdb = solv(a.v.nextSol, li_DrainBody);
  Vsb = solv(a.v.nextSol, li_SourceBody);
  Qtotal = (X && itrnqsMod) ? solv(a.v.nextSol, li_Charge) : 0.0;
  Vddp = Vd - Vdp;
  Vssp = Vs - Vsp;
  Vdbb = Vdb - Vb;
  Vdbbp = Vdb - Vbp;
  Vsbb = Vsb - Vb;
  Vsbbp = Vsb - Vbp;
  Vbpb = Vbp - Vb;
  vds = C.dtype * (Vdp - Vsp);
  vgs = C.dtype * (Vgp - Vsp);
  vbs = C.dtype * (Vbp - Vsp);
  vges = C.dtype * (Vge - Vsp);
  vgms = C.dtype * (Vgm - Vsp);
  vdbs = C.dtype * (Vdb - Vsp);
  vsbs = C.dtype * (Vsb - Vsp);
  vses = C.dtype * (Vs - Vsp);
  vdes = C.dtype * (Vd - Vsp);
  qdef = C.dtype * (Qtotal);
  vbd = vbs - vds;
  vdbd = vdbs - vds;
  vgd = vgs - vds;
  vged = vges - vds;
  vgmd = vgms - vds;
  vgmb = vgms - vbs;
  vdbd = vdbs - vds;
  vbs_jct = (!irbodyMod) ? vbs : vsbs;
  vbd_jct = (!irbodyMod) ? vbd : vdbd;
  Vgegp = Vge - Vgp;
  Vgegm = Vge - Vgm;
  Vgmgp = Vgm - Vgp;
  origFlag = 1;
  vbd_orig = vbd;
  vbs_orig = vbs;
  vgs_orig = vgs;
  vds_orig = vds;
  vgd_orig = vgd;
  vges_orig = vges;
  vgms_orig = vgms;
  vdes_orig = vdes;
  vses_orig = vses;
  vdbs_orig = vdbs;
  vsbs_orig = vsbs;
  vdbd_orig = vdbd;
  vged_orig = vged;
  vgmd_orig = vgmd;
  vbs_jct_orig = vbs_jct;
  vbd_jct_orig = vbd_jct;
  vgmb_orig = vgmb;
  vgb_orig = vgb;        /* the previous call's vgb: vgb is updated below */
  Vgegp_orig = Vgegp;
  Vgegm_orig = Vgegm;
  Vgmgp_orig = Vgmgp;

  if (RELOAD) reload_params(dfb);
  /* initial conditions (:8052-8097) */
  if (ss.initJctFlag && !IV(OFF) && opt.voltageLimiterFlag) {
    int setv = 1;
    if (ss.inputOPFlag) {
      const double* fl = a.v.flagSol;
      const int lids[11] = {li_Drain, li_GateExt, li_Source, li_Body, li_DrainPrime,
                            li_GatePrime, li_GateMid, li_SourcePrime, li_BodyPrime,
                            li_DrainBody, li_SourceBody};
      int k, any0 = (fl == nullptr);
      for (k = 0; k < 11 && !any0; ++k) any0 = (lids[k] < 0) || fl[lids[k]] == 0;
      setv = any0;
    }
    if (setv) {
      vds = 0.1;
      vdes = 0.11;
      vses = -0.01;
      vgs = vges = vgms = C.dtype * IV(vth0) + 0.1;
      origFlag = 0;
    }
    vbs = vdbs = vsbs = 0.0;
    vbd = vbs - vds;
    vdbd = vdbs - vds;
    vgd = vgs - vds;
    vged = vges - vds;
    vgmd = vgms - vds;
  } else if ((ss.initFixFlag || ss.initJctFlag) && IV(OFF)) {
    vds = vgs = vbs = vges = vgms = 0.0;
    vds = vsbs = vdes = vses = qdef = 0.0;
  }

  if (RELOAD) reload_params(dfb);
  /* previous iteration's junction voltages (:8100-8158) */
  if (ss.newtonIter == 0) {
    newtonIterOld = 0;
    if (!ss.dcopFlag || (ss.locaEnabledFlag && ss.dcopFlag)) {
      const double* s = a.v.currSto;
      vbd_old = stv(s, li_store_vbd);
      vbs_old = stv(s, li_store_vbs);
      vgs_old = stv(s, li_store_vgs);
      vds_old = stv(s, li_store_vds);
      vges_old = stv(s, li_store_vges);
      vgms_old = stv(s, li_store_vgms);
      vdes_old = stv(s, li_store_vdes);
      vses_old = stv(s, li_store_vses);
      vdbs_old = stv(s, li_store_vdbs);
      vsbs_old = stv(s, li_store_vsbs);
      vdbd_old = stv(s, li_store_vdbd);
      vged_old = stv(s, li_store_vged);
      vgmd_old = stv(s, li_store_vgmd);
      von_local = stv(s, li_store_von);
    } else {
      vbd_old = vbd;
      vbs_old = vbs;
      vgs_old = vgs;
      vds_old = vds;
      vges_old = vges;
      vgms_old = vgms;
      vdes_old = vdes;
      vses_old = vses;
      vdbs_old = vdbs;
      vsbs_old = vsbs;
      vdbd_old = vdbd;
      vged_old = vged;
      vgmd_old = vgmd;
      von_local = 0.0;
    }
  } else {
    const double* s = a.v.nextSto;
    vbd_old = stv(s, li_store_vbd);
    vbs_old = stv(s, li_store_vbs);
    vgs_old = stv(s, li_store_vgs);
    vds_old = stv(s, li_store_vds);
    vges_old = stv(s, li_store_vges);
    vgms_old = stv(s, li_store_vgms);
    vdes_old = stv(s, li_store_vdes);
    vses_old = stv(s, li_store_vses);
    vdbs_old = stv(s, li_store_vdbs);
    vsbs_old = stv(s, li_store_vsbs);
    vdbd_old = stv(s, li_store_vdbd);
    vged_old = stv(s, li_store_vged);
    vgmd_old = stv(s, li_store_vgmd);
    von_local = stv(s, li_store_von);
  }
  vgd_old = vgs_old - vds_old;
  (void)vged_old; (void)vgmd_old; (void)vges_old; (void)vgms_old; (void)vdbs_old;
  (void)vsbs_old; (void)vdbd_old; (void)vbd_old;

  if (RELOAD) reload_params(dfb);
  /* voltage limiting (:8165-8359) */
  if (opt.voltageLimiterFlag && !(ss.initFixFlag && IV(OFF))) {
    if (ss.newtonIter >= 0 && !(ss.initJctFlag)) {
      if (vds_old >= 0.0) {
        vgs = fetlim(vgs, vgs_old, von_local);
        vds = vgs - vgd;
        vds = limvds(vds, vds_old);
        vgd = vgs - vds;
        if (X) {
          if (irgateMod == 3) {
            vges = fetlim(vges, vges_old, von_local);
            vgms = fetlim(vgms, vgms_old, von_local);
          } else if ((irgateMod == 1) || (irgateMod == 2)) {
            vges = fetlim(vges, vges_old, von_local);
            vged = vges - vds;
          }
          if (rdsMod) {
            vdes = limvds(vdes, vdes_old);
            vses = -limvds(-vses, -vses_old);
          }
        }
      } else {
        vgd = fetlim(vgd, vgd_old, von_local);
        vds = vgs - vgd;
        vds = -limvds(-vds, -vds_old);
        vgs = vgd + vds;
        if (X) {
          if (irgateMod == 3) {
            vged = fetlim(vged, vged_old, von_local);
            vges = vged + vds;
            vgmd = fetlim(vgmd, vgmd_old, von_local);
            vgms = vgmd + vds;
          }
          if ((irgateMod == 1) || (irgateMod == 2)) {
            vged = fetlim(vged, vged_old, von_local);
            vges = vged + vds;
          }
          if (rdsMod) {
            vdes = -limvds(-vdes, -vdes_old);
            vses = limvds(vses, vses_old);
          }
        }
      }
      if (vds >= 0.0) {
        vbs = pnjlim(vbs, vbs_old, kVt0, C.vcrit, &Check);
        vbd = vbs - vds;
        if (X && irbodyMod) {
          int Check1, Check2;
          vdbs = pnjlim(vdbs, vdbs_old, kVt0, C.vcrit, &Check1);
          vdbd = vdbs - vds;
          vsbs = pnjlim(vsbs, vsbs_old, kVt0, C.vcrit, &Check2);
          if ((Check1 != 0) || (Check2 != 0)) Check = 1;
        }
      } else {
        vbd = pnjlim(vbd, vbd_old, kVt0, C.vcrit, &Check);
        vbs = vbd + vds;
        if (X && irbodyMod) {
          int Check1, Check2;
          vdbd = pnjlim(vdbd, vdbd_old, kVt0, C.vcrit, &Check1);
          vdbs = vdbd + vds;
          const double vsbd_old = vsbs_old - vds_old;
          double vsbd = vsbs - vds;
          vsbd = pnjlim(vsbd, vsbd_old, kVt0, C.vcrit, &Check2);
          vsbs = vsbd + vds;
          if ((Check1 != 0) || (Check2 != 0)) Check = 1;
        }
      }
    }
    if (Check == 1) limitedFlag = 1;
    {
      const double machprec = 2.220446049250313e-16;
      if (fabs(vbd_orig - vbd) > machprec || fabs(vbs_orig - vbs) > machprec ||
          fabs(vgs_orig - vgs) > machprec || fabs(vds_orig - vds) > machprec ||
          fabs(vgd_orig - vgd) > machprec || fabs(vges_orig - vges) > machprec ||
          fabs(vgms_orig - vgms) > machprec || fabs(vdes_orig - vdes) > machprec ||
          fabs(vses_orig - vses) > machprec || fabs(vdbs_orig - vdbs) > machprec ||
          fabs(vsbs_orig - vsbs) > machprec || fabs(vdbd_orig - vdbd) > machprec ||
          fabs(vbs_jct_orig - vbs_jct) > machprec || fabs(vbd_jct_orig - vbd_jct) > machprec ||
          fabs(vgmb_orig - vgmb) > machprec || fabs(vgb_orig - vgb) > machprec ||
          fabs(vged_orig - vged) > machprec || fabs(vgmd_orig - vgmd) > machprec)
        origFlag = 0;
    }
  }
  if (ss.newtonIter != 0 && ss.newtonIter != newtonIterOld) newtonIterOld = ss.newtonIter;
  vbd = vbs - vds;
  vgd = vgs - vds;
  vgb = vgs - vbs;
  vged = vges - vds;
  vgmd = vgms - vds;
  vgmb = vgms - vbs;
  vdbd = vdbs - vds;
  vbs_jct = (!irbodyMod) ? vbs : vsbs;
  vbd_jct = (!irbodyMod) ? vbd : vdbd;

  if (RELOAD) reload_params(dfb);
  /* source/drain junction diodes (:8425-8769) */
  Nvtms = C.vtm * C.SjctEmissionCoeff;
  if ((IV(Aseff) <= 0.0) && (IV(Pseff) <= 0.0)) SourceSatCurrent = 1.0e-14;
  else SourceSatCurrent = IV(Aseff) * C.SjctTempSatCurDensity + IV(Pseff) * C.SjctSidewallTempSatCurDensity
                          + PP(weffCJ) * nf * C.SjctGateSidewallTempSatCurDensity;
  if (SourceSatCurrent <= 0.0) {
    gbs = gmin;
    cbs = gbs * vbs_jct;
  } else {
    switch ((int)C.dioMod) {
      case 0:
        evbs = exp(vbs_jct / Nvtms);
        T1 = C.xjbvs * exp(-(C.bvs + vbs_jct) / Nvtms);
        gbs = SourceSatCurrent * (evbs + T1) / Nvtms + gmin;
        cbs = SourceSatCurrent * (evbs + IV(XExpBVS) - T1 - 1.0) + gmin * vbs_jct;
        break;
      case 1:
        T2 = vbs_jct / Nvtms;
        if (T2 < -kExpThreshold) {
          gbs = gmin;
          cbs = SourceSatCurrent * (kMinExp - 1.0) + gmin * vbs_jct;
        } else if (vbs_jct <= IV(vjsmFwd)) {
          evbs = exp(T2);
          gbs = SourceSatCurrent * evbs / Nvtms + gmin;
          cbs = SourceSatCurrent * (evbs - 1.0) + gmin * vbs_jct;
        } else {
          T0 = IV(IVjsmFwd) / Nvtms;
          gbs = T0 + gmin;
          cbs = IV(IVjsmFwd) - SourceSatCurrent + T0 * (vbs_jct - IV(vjsmFwd)) + gmin * vbs_jct;
        }
        break;
      case 2:
        if (vbs_jct < IV(vjsmRev)) {
          T0 = vbs_jct / Nvtms;
          if (T0 < -kExpThreshold) {
            evbs = kMinExp;
            devbs_dvb = 0.0;
          } else {
            evbs = exp(T0);
            devbs_dvb = evbs / Nvtms;
          }
          T1 = evbs - 1.0;
          T2 = IV(IVjsmRev) + IV(SslpRev) * (vbs_jct - IV(vjsmRev));
          gbs = devbs_dvb * T2 + T1 * IV(SslpRev) + gmin;
          cbs = T1 * T2 + gmin * vbs_jct;
        } else if (vbs_jct <= IV(vjsmFwd)) {
          T0 = vbs_jct / Nvtms;
          if (T0 < -kExpThreshold) {
            evbs = kMinExp;
            devbs_dvb = 0.0;
          } else {
            evbs = exp(T0);
            devbs_dvb = evbs / Nvtms;
          }
          T1 = (C.bvs + vbs_jct) / Nvtms;
          if (T1 > kExpThreshold) {
            T2 = kMinExp;
            T3 = 0.0;
          } else {
            T2 = exp(-T1);
            T3 = -T2 / Nvtms;
          }
          gbs = SourceSatCurrent * (devbs_dvb - C.xjbvs * T3) + gmin;
          cbs = SourceSatCurrent * (evbs + IV(XExpBVS) - 1.0 - C.xjbvs * T2) + gmin * vbs_jct;
        } else {
          gbs = IV(SslpFwd) + gmin;
          cbs = IV(IVjsmFwd) + IV(SslpFwd) * (vbs_jct - IV(vjsmFwd)) + gmin * vbs_jct;
        }
        break;
      default: gbs = cbs = 0.0; break;   // dioMod is checked at setup
    }
  }
  Nvtmd = C.vtm * C.DjctEmissionCoeff;
  if ((IV(Adeff) <= 0.0) && (IV(Pdeff) <= 0.0)) DrainSatCurrent = 1.0e-14;
  else DrainSatCurrent = IV(Adeff) * C.DjctTempSatCurDensity + IV(Pdeff) * C.DjctSidewallTempSatCurDensity
                         + PP(weffCJ) * nf * C.DjctGateSidewallTempSatCurDensity;
  if (DrainSatCurrent <= 0.0) {
    gbd = gmin;
    cbd = gbd * vbd_jct;
  } else {
    switch ((int)C.dioMod) {
      case 0:
        evbd = exp(vbd_jct / Nvtmd);
        T1 = C.xjbvd * exp(-(C.bvd + vbd_jct) / Nvtmd);
        gbd = DrainSatCurrent * (evbd + T1) / Nvtmd + gmin;
        cbd = DrainSatCurrent * (evbd + IV(XExpBVD) - T1 - 1.0) + gmin * vbd_jct;
        break;
      case 1:
        T2 = vbd_jct / Nvtmd;
        if (T2 < -kExpThreshold) {
          gbd = gmin;
          cbd = DrainSatCurrent * (kMinExp - 1.0) + gmin * vbd_jct;
        } else if (vbd_jct <= IV(vjdmFwd)) {
          evbd = exp(T2);
          gbd = DrainSatCurrent * evbd / Nvtmd + gmin;
          cbd = DrainSatCurrent * (evbd - 1.0) + gmin * vbd_jct;
        } else {
          T0 = IV(IVjdmFwd) / Nvtmd;
          gbd = T0 + gmin;
          cbd = IV(IVjdmFwd) - DrainSatCurrent + T0 * (vbd_jct - IV(vjdmFwd)) + gmin * vbd_jct;
        }
        break;
      case 2:
        if (vbd_jct < IV(vjdmRev)) {
          T0 = vbd_jct / Nvtmd;
          if (T0 < -kExpThreshold) {
            evbd = kMinExp;
            devbd_dvb = 0.0;
          } else {
            evbd = exp(T0);
            devbd_dvb = evbd / Nvtmd;
          }
          T1 = evbd - 1.0;
          T2 = IV(IVjdmRev) + IV(DslpRev) * (vbd_jct - IV(vjdmRev));
          gbd = devbd_dvb * T2 + T1 * IV(DslpRev) + gmin;
          cbd = T1 * T2 + gmin * vbd_jct;
        } else if (vbd_jct <= IV(vjdmFwd)) {
          T0 = vbd_jct / Nvtmd;
          if (T0 < -kExpThreshold) {
            evbd = kMinExp;
            devbd_dvb = 0.0;
          } else {
            evbd = exp(T0);
            devbd_dvb = evbd / Nvtmd;
          }
          T1 = (C.bvd + vbd_jct) / Nvtmd;
          if (T1 > kExpThreshold) {
            T2 = kMinExp;
            T3 = 0.0;
          } else {
            T2 = exp(-T1);
            T3 = -T2 / Nvtmd;
          }
          gbd = DrainSatCurrent * (devbd_dvb - C.xjbvd * T3) + gmin;
          cbd = DrainSatCurrent * (evbd + IV(XExpBVD) - 1.0 - C.xjbvd * T2) + gmin * vbd_jct;
        } else {
          gbd = IV(DslpFwd) + gmin;
          cbd = IV(IVjdmFwd) + IV(DslpFwd) * (vbd_jct - IV(vjdmFwd)) + gmin * vbd_jct;
        }
        break;
      default: gbd = cbd = 0.0; break;
    }
  }

  if (RELOAD) reload_params(dfb);
  /* trap-assisted tunneling and recombination (:8655-8769) */
  Nvtmrssws = C.vtm0 * C.njtsswstemp;
  Nvtmrsswgs = C.vtm0 * C.njtsswgstemp;
  Nvtmrss = C.vtm0 * C.njtsstemp;
  Nvtmrsswd = C.vtm0 * C.njtsswdtemp;
  Nvtmrsswgd = C.vtm0 * C.njtsswgdtemp;
  Nvtmrsd = C.vtm0 * C.njtsdtemp;
  if ((C.vtss - vbs_jct) < (C.vtss * 1e-3)) {
    T9 = 1.0e3;
    T0 = -vbs_jct / Nvtmrss * T9;
    dexp3(T0, T1, T10);
    dT1_dVb = T10 / Nvtmrss * T9;
  } else {
    T9 = 1.0 / (C.vtss - vbs_jct);
    T0 = -vbs_jct / Nvtmrss * C.vtss * T9;
    dT0_dVb = C.vtss / Nvtmrss * (T9 + vbs_jct * T9 * T9);
    dexp3(T0, T1, T10);
    dT1_dVb = T10 * dT0_dVb;
  }
  if ((C.vtsd - vbd_jct) < (C.vtsd * 1e-3)) {
    T9 = 1.0e3;
    T0 = -vbd_jct / Nvtmrsd * T9;
    dexp3(T0, T2, T10);
    dT2_dVb = T10 / Nvtmrsd * T9;
  } else {
    T9 = 1.0 / (C.vtsd - vbd_jct);
    T0 = -vbd_jct / Nvtmrsd * C.vtsd * T9;
    dT0_dVb = C.vtsd / Nvtmrsd * (T9 + vbd_jct * T9 * T9);
    dexp3(T0, T2, T10);
    dT2_dVb = T10 * dT0_dVb;
  }
  if ((C.vtssws - vbs_jct) < (C.vtssws * 1e-3)) {
    T9 = 1.0e3;
    T0 = -vbs_jct / Nvtmrssws * T9;
    dexp3(T0, T3, T10);
    dT3_dVb = T10 / Nvtmrssws * T9;
  } else {
    T9 = 1.0 / (C.vtssws - vbs_jct);
    T0 = -vbs_jct / Nvtmrssws * C.vtssws * T9;
    dT0_dVb = C.vtssws / Nvtmrssws * (T9 + vbs_jct * T9 * T9);
    dexp3(T0, T3, T10);
    dT3_dVb = T10 * dT0_dVb;
  }
  if ((C.vtsswd - vbd_jct) < (C.vtsswd * 1e-3)) {
    T9 = 1.0e3;
    T0 = -vbd_jct / Nvtmrsswd * T9;
    dexp3(T0, T4, T10);
    dT4_dVb = T10 / Nvtmrsswd * T9;
  } else {
    T9 = 1.0 / (C.vtsswd - vbd_jct);
    T0 = -vbd_jct / Nvtmrsswd * C.vtsswd * T9;
    dT0_dVb = C.vtsswd / Nvtmrsswd * (T9 + vbd_jct * T9 * T9);
    dexp3(T0, T4, T10);
    dT4_dVb = T10 * dT0_dVb;
  }
  if ((C.vtsswgs - vbs_jct) < (C.vtsswgs * 1e-3)) {
    T9 = 1.0e3;
    T0 = -vbs_jct / Nvtmrsswgs * T9;
    dexp3(T0, T5, T10);
    dT5_dVb = T10 / Nvtmrsswgs * T9;
  } else {
    T9 = 1.0 / (C.vtsswgs - vbs_jct);
    T0 = -vbs_jct / Nvtmrsswgs * C.vtsswgs * T9;
    dT0_dVb = C.vtsswgs / Nvtmrsswgs * (T9 + vbs_jct * T9 * T9);
    dexp3(T0, T5, T10);
    dT5_dVb = T10 * dT0_dVb;
  }
  if ((C.vtsswgd - vbd_jct) < (C.vtsswgd * 1e-3)) {
    T9 = 1.0e3;
    T0 = -vbd_jct / Nvtmrsswgd * T9;
    dexp3(T0, T6, T10);
    dT6_dVb = T10 / Nvtmrsswgd * T9;
  } else {
    T9 = 1.0 / (C.vtsswgd - vbd_jct);
    T0 = -vbd_jct / Nvtmrsswgd * C.vtsswgd * T9;
    dT0_dVb = C.vtsswgd / Nvtmrsswgd * (T9 + vbd_jct * T9 * T9);
    dexp3(T0, T6, T10);
    dT6_dVb = T10 * dT0_dVb;
  }
  gbs += IV(SjctTempRevSatCur) * dT1_dVb + IV(SswTempRevSatCur) * dT3_dVb
            + IV(SswgTempRevSatCur) * dT5_dVb;
  cbs -= IV(SjctTempRevSatCur) * (T1 - 1.0) + IV(SswTempRevSatCur) * (T3 - 1.0)
            + IV(SswgTempRevSatCur) * (T5 - 1.0);
  gbd += IV(DjctTempRevSatCur) * dT2_dVb + IV(DswTempRevSatCur) * dT4_dVb
            + IV(DswgTempRevSatCur) * dT6_dVb;
  cbd -= IV(DjctTempRevSatCur) * (T2 - 1.0) + IV(DswTempRevSatCur) * (T4 - 1.0)
            + IV(DswgTempRevSatCur) * (T6 - 1.0);

  if (RELOAD) reload_params(dfb);
  /* mode (:8773-8788) */
  if (vds >= 0.0) {
    mode = 1;
    Vds = vds;
    Vgs = vgs;
    Vbs = vbs;
    Vdb_ = vds - vbs;
  } else {
    mode = -1;
    Vds = -vds;
    Vgs = vgd;
    Vbs = vbd;
    Vdb_ = -vbs;
  }
  (void)Vdb_;
  if (X && mtrlMod) {
    epsrox = 3.9;
    toxe = XP(eot);
    epssub = kEps0 * XP(epsrsub);
  } else {
    epsrox = C.epsrox;
    toxe = C.toxe;
    epssub = kEpsSi;
  }
  if (RELOAD) reload_params(dfb);
  /* MOSFET homotopy (:8818-8838), one gain-scale block, no staggering */
  if (ss.artParameterFlag) {
    const double alpha = ss.gainScale;
    const double vgstConst = 4.5, vdsScaleMin = 0.3;
    Vds = Vds * (ss.nltermScale * (1.0 - vdsScaleMin) + vdsScaleMin);
    Vgs = (alpha) * Vgs + (1.0 - alpha) * vgstConst;
  }

  if (RELOAD) reload_params(dfb);
  /* Vbseff (:8849-8868) */
  T0 = Vbs - IV(vbsc) - 0.001;
  T1 = sqrt(T0 * T0 - 0.004 * IV(vbsc));
  if (T0 >= 0.0) {
    Vbseff = IV(vbsc) + 0.5 * (T0 + T1);
    dVbseff_dVb = 0.5 * (1.0 + T0 / T1);
  } else {
    T2 = -0.002 / (T1 - T0);
    Vbseff = IV(vbsc) * (1.0 + T2);
    dVbseff_dVb = T2 * IV(vbsc) / T1;
  }
  T9 = 0.95 * PP(phi);
  T0 = T9 - Vbseff - 0.001;
  T1 = sqrt(T0 * T0 + 0.004 * T9);
  Vbseff = T9 - 0.5 * (T0 + T1);
  dVbseff_dVb *= 0.5 * (1.0 + T0 / T1);
  Phis = PP(phi) - Vbseff;
  dPhis_dVb = -1.0;
  sqrtPhis = sqrt(Phis);
  dsqrtPhis_dVb = -0.5 / sqrtPhis;
  Xdep = PP(Xdep0) * sqrtPhis / PP(sqrtPhi);
  dXdep_dVb = (PP(Xdep0) / PP(sqrtPhi)) * dsqrtPhis_dVb;
  Leff = PP(leff);
  Vtm = C.vtm;
  Vtm0 = C.vtm0;

  if (RELOAD) reload_params(dfb);
  /* threshold voltage (:8883-9052) */
  T3 = sqrt(Xdep);
  V0 = PP(vbi) - PP(phi);
  T0 = PP(dvt2) * Vbseff;
  if (T0 >= -0.5) {
    T1 = 1.0 + T0;
    T2 = PP(dvt2);
  } else {
    T4 = 1.0 / (3.0 + 8.0 * T0);
    T1 = (1.0 + 3.0 * T0) * T4;
    T2 = PP(dvt2) * T4 * T4;
  }
  lt1 = C.factor1 * T3 * T1;
  dlt1_dVb = C.factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
  T0 = PP(dvt2w) * Vbseff;
  if (T0 >= -0.5) {
    T1 = 1.0 + T0;
    T2 = PP(dvt2w);
  } else {
    T4 = 1.0 / (3.0 + 8.0 * T0);
    T1 = (1.0 + 3.0 * T0) * T4;
    T2 = PP(dvt2w) * T4 * T4;
  }
  ltw = C.factor1 * T3 * T1;
  dltw_dVb = C.factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
  T0 = PP(dvt1) * Leff / lt1;
  if (T0 < kExpThreshold) {
    T1 = exp(T0);
    T2 = T1 - 1.0;
    T3 = T2 * T2;
    T4 = T3 + 2.0 * T1 * kMinExp;
    Theta0 = T1 / T4;
    dT1_dVb = -T0 * T1 * dlt1_dVb / lt1;
    dTheta0_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + kMinExp)) / T4 / T4;
  } else {
    Theta0 = 1.0 / (kMaxExp - 2.0);
    dTheta0_dVb = 0.0;
  }
  thetavth = PP(dvt0) * Theta0;
  Delt_vth = thetavth * V0;
  dDelt_vth_dVb = PP(dvt0) * dTheta0_dVb * V0;
  T0 = PP(dvt1w) * PP(weff) * Leff / ltw;
  if (T0 < kExpThreshold) {
    T1 = exp(T0);
    T2 = T1 - 1.0;
    T3 = T2 * T2;
    T4 = T3 + 2.0 * T1 * kMinExp;
    T5 = T1 / T4;
    dT1_dVb = -T0 * T1 * dltw_dVb / ltw;
    dT5_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + kMinExp)) / T4 / T4;
  } else {
    T5 = 1.0 / (kMaxExp - 2.0);
    dT5_dVb = 0.0;
  }
  T0 = PP(dvt0w) * T5;
  T2 = T0 * V0;
  dT2_dVb = PP(dvt0w) * dT5_dVb * V0;
  TempRatio = IV(temp) / C.tnom - 1.0;
  T0 = sqrt(1.0 + PP(lpe0) / Leff);
  T1 = PP(k1ox) * (T0 - 1.0) * PP(sqrtPhi) + (PP(kt1) + PP(kt1l) / Leff + PP(kt2) * Vbseff) * TempRatio;
  Vth_NarrowW = toxe * PP(phi) / (PP(weff) + PP(w0));
  T3 = IV(eta0) + PP(etab) * Vbseff;
  if (T3 < 1.0e-4) {
    T9 = 1.0 / (3.0 - 2.0e4 * T3);
    T3 = (2.0e-4 - T3) * T9;
    T4 = T9 * T9;
  } else {
    T4 = 1.0;
  }
  dDIBL_Sft_dVd = T3 * PP(theta0vb0);
  DIBL_Sft = dDIBL_Sft_dVd * Vds;
  Lpe_Vb = sqrt(1.0 + PP(lpeb) / Leff);
  Vth = C.dtype * IV(vth0) + (PP(k1ox) * sqrtPhis - PP(k1) * PP(sqrtPhi)) * Lpe_Vb
        - IV(k2ox) * Vbseff - Delt_vth - T2 + (PP(k3) + PP(k3b) * Vbseff) * Vth_NarrowW + T1 - DIBL_Sft;
  dVth_dVb = Lpe_Vb * PP(k1ox) * dsqrtPhis_dVb - IV(k2ox) - dDelt_vth_dVb - dT2_dVb
             + PP(k3b) * Vth_NarrowW - PP(etab) * Vds * PP(theta0vb0) * T4 + PP(kt2) * TempRatio;
  dVth_dVd = -dDIBL_Sft_dVd;
  tmp1 = epssub / Xdep;
  nstar = C.vtm / kQ * (C.coxe + tmp1 + PP(cit));
  tmp2 = PP(nfactor) * tmp1;
  tmp3 = PP(cdsc) + PP(cdscb) * Vbseff + PP(cdscd) * Vds;
  tmp4 = (tmp2 + tmp3 * Theta0 + PP(cit)) / C.coxe;
  if (tmp4 >= -0.5) {
    n = 1.0 + tmp4;
    dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + PP(cdscb) * Theta0) / C.coxe;
    dn_dVd = PP(cdscd) * Theta0 / C.coxe;
  } else {
    T0 = 1.0 / (3.0 + 8.0 * tmp4);
    n = (1.0 + 3.0 * tmp4) * T0;
    T0 *= T0;
    dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + PP(cdscb) * Theta0) / C.coxe * T0;
    dn_dVd = PP(cdscd) * Theta0 / C.coxe * T0;
  }
  if (PP(dvtp0) > 0.0) {
    T0 = -PP(dvtp1) * Vds;
    if (T0 < -kExpThreshold) {
      T2 = kMinExp;
      dT2_dVd = 0.0;
    } else {
      T2 = exp(T0);
      dT2_dVd = -PP(dvtp1) * T2;
    }
    T3 = Leff + PP(dvtp0) * (1.0 + T2);
    dT3_dVd = PP(dvtp0) * dT2_dVd;
    if (C.tempMod < 2) {
      T4 = Vtm * log(Leff / T3);
      dT4_dVd = -Vtm * dT3_dVd / T3;
    } else {
      T4 = C.vtm0 * log(Leff / T3);
      dT4_dVd = -C.vtm0 * dT3_dVd / T3;
    }
    dDITS_Sft_dVd = dn_dVd * T4 + n * dT4_dVd;
    dDITS_Sft_dVb = T4 * dn_dVb;
    Vth -= n * T4;
    dVth_dVd -= dDITS_Sft_dVd;
    dVth_dVb -= dDITS_Sft_dVb;
  }
  von = Vth;

  if (RELOAD) reload_params(dfb);
  /* poly gate depletion (:9055-9080) */
  T0 = IV(vfb) + PP(phi);
  if (X && mtrlMod) T1 = XP(epsrgate) * kEps0;
  else T1 = kEpsSi;
  polyDepletion(T0, PP(ngate), T1, C.coxe, vgs, vgs_eff, dvgs_eff_dvg);
  polyDepletion(T0, PP(ngate), T1, C.coxe, vgd, vgd_eff, dvgd_eff_dvg);
  if (mode > 0) {
    Vgs_eff = vgs_eff;
    dVgs_eff_dVg = dvgs_eff_dvg;
  } else {
    Vgs_eff = vgd_eff;
    dVgs_eff_dVg = dvgd_eff_dvg;
  }
  Vgst = Vgs_eff - Vth;

  if (RELOAD) reload_params(dfb);
  /* effective Vgst (:9085-9150) */
  T0 = n * Vtm;
  T1 = PP(mstar) * Vgst;
  T2 = T1 / T0;
  if (T2 > kExpThreshold) {
    T10 = T1;
    dT10_dVg = PP(mstar) * dVgs_eff_dVg;
    dT10_dVd = -dVth_dVd * PP(mstar);
    dT10_dVb = -dVth_dVb * PP(mstar);
  } else if (T2 < -kExpThreshold) {
    T10 = Vtm * log(1.0 + kMinExp);
    dT10_dVg = 0.0;
    dT10_dVd = T10 * dn_dVd;
    dT10_dVb = T10 * dn_dVb;
    T10 *= n;
  } else {
    ExpVgst = exp(T2);
    T3 = Vtm * log(1.0 + ExpVgst);
    T10 = n * T3;
    dT10_dVg = PP(mstar) * ExpVgst / (1.0 + ExpVgst);
    dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
    dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
    dT10_dVg *= dVgs_eff_dVg;
  }
  T1 = PP(voffcbn) - (1.0 - PP(mstar)) * Vgst;
  T2 = T1 / T0;
  if (T2 < -kExpThreshold) {
    T3 = C.coxe * kMinExp / PP(cdep0);
    T9 = PP(mstar) + T3 * n;
    dT9_dVg = 0.0;
    dT9_dVd = dn_dVd * T3;
    dT9_dVb = dn_dVb * T3;
  } else if (T2 > kExpThreshold) {
    T3 = C.coxe * kMaxExp / PP(cdep0);
    T9 = PP(mstar) + T3 * n;
    dT9_dVg = 0.0;
    dT9_dVd = dn_dVd * T3;
    dT9_dVb = dn_dVb * T3;
  } else {
    ExpVgst = exp(T2);
    T3 = C.coxe / PP(cdep0);
    T4 = T3 * ExpVgst;
    T5 = T1 * T4 / T0;
    T9 = PP(mstar) + n * T4;
    dT9_dVg = T3 * (PP(mstar) - 1.0) * ExpVgst / Vtm;
    dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
    dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
    dT9_dVg *= dVgs_eff_dVg;
  }
  Vgsteff = T10 / T9;
  T11 = T9 * T9;
  dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
  dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
  dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;

  if (RELOAD) reload_params(dfb);
  /* effective channel geometry and series resistance (:9153-9198) */
  T9 = sqrtPhis - PP(sqrtPhi);
  Weff = PP(weff) - 2.0 * (PP(dwg) * Vgsteff + PP(dwb) * T9);
  dWeff_dVg = -2.0 * PP(dwg);
  dWeff_dVb = -2.0 * PP(dwb) * dsqrtPhis_dVb;
  if (Weff < 2.0e-8) {
    T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
    Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
    T0 *= T0 * 4.0e-16;
    dWeff_dVg *= T0;
    dWeff_dVb *= T0;
  }
  if (C.rdsMod == 1) {
    Rds = dRds_dVg = dRds_dVb = 0.0;
  } else {
    T0 = 1.0 + PP(prwg) * Vgsteff;
    dT0_dVg = -PP(prwg) / T0 / T0;
    T1 = PP(prwb) * T9;
    dT1_dVb = PP(prwb) * dsqrtPhis_dVb;
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    dT3_dVg = 1.0 + T2 / (T3 - T2);
    dT3_dVb = dT3_dVg * dT1_dVb;
    dT3_dVg *= dT0_dVg;
    T4 = PP(rds0) * 0.5;
    Rds = PP(rdswmin) + T3 * T4;
    dRds_dVg = T4 * dT3_dVg;
    dRds_dVb = T4 * dT3_dVb;
    if (Rds > 0.0) grdsw = 1.0 / Rds;
    else grdsw = 0.0;
  }

  if (RELOAD) reload_params(dfb);
  /* bulk charge effect (:9201-9258) */
  T9 = 0.5 * PP(k1ox) * Lpe_Vb / sqrtPhis;
  T1 = T9 + IV(k2ox) - PP(k3b) * Vth_NarrowW;
  dT1_dVb = -T9 / sqrtPhis * dsqrtPhis_dVb;
  T9 = sqrt(PP(xj) * Xdep);
  tmp1 = Leff + 2.0 * T9;
  T5 = Leff / tmp1;
  tmp2 = PP(a0) * T5;
  tmp3 = PP(weff) + PP(b1);
  tmp4 = PP(b0) / tmp3;
  T2 = tmp2 + tmp4;
  dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
  T6 = T5 * T5;
  T7 = T5 * T6;
  Abulk0 = 1.0 + T1 * T2;
  dAbulk0_dVb = T1 * tmp2 * dT2_dVb + T2 * dT1_dVb;
  T8 = PP(ags) * PP(a0) * T7;
  dAbulk_dVg = -T1 * T8;
  Abulk = Abulk0 + dAbulk_dVg * Vgsteff;
  dAbulk_dVb = dAbulk0_dVb - T8 * Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb);
  if (Abulk0 < 0.1) {
    T9 = 1.0 / (3.0 - 20.0 * Abulk0);
    Abulk0 = (0.2 - Abulk0) * T9;
    dAbulk0_dVb *= T9 * T9;
  }
  if (Abulk < 0.1) {
    T9 = 1.0 / (3.0 - 20.0 * Abulk);
    Abulk = (0.2 - Abulk) * T9;
    T10 = T9 * T9;
    dAbulk_dVb *= T10;
    dAbulk_dVg *= T10;
  }
  T2 = PP(keta) * Vbseff;
  if (T2 >= -0.9) {
    T0 = 1.0 / (1.0 + T2);
    dT0_dVb = -PP(keta) * T0 * T0;
  } else {
    T1 = 1.0 / (0.8 + T2);
    T0 = (17.0 + 20.0 * T2) * T1;
    dT0_dVb = -PP(keta) * T1 * T1;
  }
  dAbulk_dVg *= T0;
  dAbulk_dVb = dAbulk_dVb * T0 + Abulk * dT0_dVb;
  dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVb;
  Abulk *= T0;
  Abulk0 *= T0;

  if (RELOAD) reload_params(dfb);
  /* mobility (:9261-9341) */
  if (X && mtrlMod) T14 = 2.0 * C.dtype * (XP(phig) - XP(easub) - 0.5 * XP(Eg0) + 0.45);
  else T14 = 0.0;
  if (C.mobMod == 0) {
    T0 = Vgsteff + Vth + Vth - T14;
    T2 = PP(ua) + PP(uc) * Vbseff;
    T3 = T0 / toxe;
    T12 = sqrt(Vth * Vth + 0.0001);
    T9 = 1.0 / (Vgsteff + 2 * T12);
    T10 = T9 * toxe;
    T8 = PP(ud) * T10 * T10 * Vth;
    T6 = T8 * Vth;
    T5 = T3 * (T2 + PP(ub) * T3) + T6;
    T7 = -2.0 * T6 * T9;
    T11 = T7 * Vth / T12;
    dDenomi_dVg = (T2 + 2.0 * PP(ub) * T3) / toxe;
    T13 = 2.0 * (dDenomi_dVg + T11 + T8);
    dDenomi_dVd = T13 * dVth_dVd;
    dDenomi_dVb = T13 * dVth_dVb + PP(uc) * T3;
    dDenomi_dVg += T7;
  } else if (C.mobMod == 1) {
    T0 = Vgsteff + Vth + Vth - T14;
    T2 = 1.0 + PP(uc) * Vbseff;
    T3 = T0 / toxe;
    T4 = T3 * (PP(ua) + PP(ub) * T3);
    T12 = sqrt(Vth * Vth + 0.0001);
    T9 = 1.0 / (Vgsteff + 2 * T12);
    T10 = T9 * toxe;
    T8 = PP(ud) * T10 * T10 * Vth;
    T6 = T8 * Vth;
    T5 = T4 * T2 + T6;
    T7 = -2.0 * T6 * T9;
    T11 = T7 * Vth / T12;
    dDenomi_dVg = (PP(ua) + 2.0 * PP(ub) * T3) * T2 / toxe;
    T13 = 2.0 * (dDenomi_dVg + T11 + T8);
    dDenomi_dVd = T13 * dVth_dVd;
    dDenomi_dVb = T13 * dVth_dVb + PP(uc) * T4;
    dDenomi_dVg += T7;
  } else {
    T0 = (Vgsteff + IV(vtfbphi1)) / toxe;
    T1 = exp(PP(eu) * log(T0));
    dT1_dVg = T1 * PP(eu) / T0 / toxe;
    T2 = PP(ua) + PP(uc) * Vbseff;
    T3 = T0 / toxe;
    T12 = sqrt(Vth * Vth + 0.0001);
    T9 = 1.0 / (Vgsteff + 2 * T12);
    T10 = T9 * toxe;
    T8 = PP(ud) * T10 * T10 * Vth;
    T6 = T8 * Vth;
    T5 = T1 * T2 + T6;
    T7 = -2.0 * T6 * T9;
    T11 = T7 * Vth / T12;
    dDenomi_dVg = T2 * dT1_dVg + T7;
    T13 = 2.0 * (T11 + T8);
    dDenomi_dVd = T13 * dVth_dVd;
    dDenomi_dVb = T13 * dVth_dVb + T1 * PP(uc);
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
  ueff = IV(u0temp) / Denomi;
  T9 = -ueff / Denomi;
  dueff_dVg = T9 * dDenomi_dVg;
  dueff_dVd = T9 * dDenomi_dVd;
  dueff_dVb = T9 * dDenomi_dVb;

  if (RELOAD) reload_params(dfb);
  /* saturation field and Vdsat (:9344-9450) */
  WVCox = Weff * IV(vsattemp) * C.coxe;
  WVCoxRds = WVCox * Rds;
  Esat = 2.0 * IV(vsattemp) / ueff;
  EsatL = Esat * Leff;
  T0 = -EsatL / ueff;
  dEsatL_dVg = T0 * dueff_dVg;
  dEsatL_dVd = T0 * dueff_dVd;
  dEsatL_dVb = T0 * dueff_dVb;
  a1 = PP(a1);
  if (a1 == 0.0) {
    Lambda = PP(a2);
    dLambda_dVg = 0.0;
  } else if (a1 > 0.0) {
    T0 = 1.0 - PP(a2);
    T1 = T0 - PP(a1) * Vgsteff - 0.0001;
    T2 = sqrt(T1 * T1 + 0.0004 * T0);
    Lambda = PP(a2) + T0 - 0.5 * (T1 + T2);
    dLambda_dVg = 0.5 * PP(a1) * (1.0 + T1 / T2);
  } else {
    T1 = PP(a2) + PP(a1) * Vgsteff - 0.0001;
    T2 = sqrt(T1 * T1 + 0.0004 * PP(a2));
    Lambda = 0.5 * (T1 + T2);
    dLambda_dVg = 0.5 * PP(a1) * (1.0 + T1 / T2);
  }
  Vgst2Vtm = Vgsteff + 2.0 * Vtm;
  if (Rds > 0) {
    tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff;
    tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff;
  } else {
    tmp2 = dWeff_dVg / Weff;
    tmp3 = dWeff_dVb / Weff;
  }
  if ((Rds == 0.0) && (Lambda == 1.0)) {
    T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
    tmp1 = 0.0;
    T1 = T0 * T0;
    T2 = Vgst2Vtm * T0;
    T3 = EsatL * Vgst2Vtm;
    Vdsat = T3 * T0;
    dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
    dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
    dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
    dVdsat_dVg = T3 * dT0_dVg + T2 * dEsatL_dVg + EsatL * T0;
    dVdsat_dVd = T3 * dT0_dVd + T2 * dEsatL_dVd;
    dVdsat_dVb = T3 * dT0_dVb + T2 * dEsatL_dVb;
  } else {
    tmp1 = dLambda_dVg / (Lambda * Lambda);
    T9 = Abulk * WVCoxRds;
    T8 = Abulk * T9;
    T7 = Vgst2Vtm * T9;
    T6 = Vgst2Vtm * WVCoxRds;
    T0 = 2.0 * Abulk * (T9 - 1.0 + 1.0 / Lambda);
    dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 + (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
    dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) + (1.0 / Lambda - 1.0) * dAbulk_dVb);
    dT0_dVd = 0.0;
    T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
    dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg
              + EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
    dT1_dVb = Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
    dT1_dVd = Abulk * dEsatL_dVd;
    T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
    dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
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
  vdsat = Vdsat;

  if (RELOAD) reload_params(dfb);
  /* effective Vds (:9453-9497) */
  T1 = Vdsat - Vds - PP(delta);
  dT1_dVg = dVdsat_dVg;
  dT1_dVd = dVdsat_dVd - 1.0;
  dT1_dVb = dVdsat_dVb;
  T2 = sqrt(T1 * T1 + 4.0 * PP(delta) * Vdsat);
  T0 = T1 / T2;
  T9 = 2.0 * PP(delta);
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
  Vdseff = Vdseff_;

  if (RELOAD) reload_params(dfb);
  /* velocity overshoot (:9500-9540) */
  if ((C.lambdaGiven) && (C.lambda > 0.0)) {
    T1 = Leff * ueff;
    T2 = PP(lambda) / T1;
    T3 = -T2 / T1 * Leff;
    dT2_dVd = T3 * dueff_dVd;
    dT2_dVg = T3 * dueff_dVg;
    dT2_dVb = T3 * dueff_dVb;
    T5 = 1.0 / (Esat * PP(litl));
    T4 = -T5 / EsatL;
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
    dEsatL_dVg += EsatL * dT10_dVg;
    dEsatL_dVd *= T10;
    dEsatL_dVd += EsatL * dT10_dVd;
    dEsatL_dVb *= T10;
    dEsatL_dVb += EsatL * dT10_dVb;
    EsatL *= T10;
  }

  if (RELOAD) reload_params(dfb);
  /* Vasat (:9543-9565) */
  tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
  T9 = WVCoxRds * Vgsteff;
  T8 = T9 / Vgst2Vtm;
  T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
  T7 = 2.0 * WVCoxRds * tmp4;
  dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff)
            - T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
  dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff
            - T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
  dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
  T9 = WVCoxRds * Abulk;
  T1 = 2.0 / Lambda - 1.0 + T9;
  dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
  dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
  Vasat = T0 / T1;
  dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
  dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
  dVasat_dVd = dT0_dVd / T1;

  if (RELOAD) reload_params(dfb);
  /* effective oxide capacitance, beta, Idl (:9568-9622) */
  tmp1 = IV(vtfbphi2);
  tmp2 = 2.0e8 * C.toxp;
  dT0_dVg = 1.0 / tmp2;
  T0 = (Vgsteff + tmp1) * dT0_dVg;
  tmp3 = exp(C.bdos * 0.7 * log(T0));
  T1 = 1.0 + tmp3;
  T2 = C.bdos * 0.7 * tmp3 / T0;
  Tcen = C.ados * 1.9e-9 / T1;
  dTcen_dVg = -Tcen * T2 * dT0_dVg / T1;
  Coxeff = epssub * C.coxp / (epssub + C.coxp * Tcen);
  dCoxeff_dVg = -Coxeff * Coxeff * dTcen_dVg / epssub;
  CoxeffWovL = Coxeff * Weff / Leff;
  beta = ueff * CoxeffWovL;
  T3 = ueff / Leff;
  dbeta_dVg = CoxeffWovL * dueff_dVg + T3 * (Weff * dCoxeff_dVg + Coxeff * dWeff_dVg);
  dbeta_dVd = CoxeffWovL * dueff_dVd;
  dbeta_dVb = CoxeffWovL * dueff_dVb + T3 * Coxeff * dWeff_dVb;
  AbovVgst2Vtm = Abulk / Vgst2Vtm;
  T0 = 1.0 - 0.5 * Vdseff_ * AbovVgst2Vtm;
  dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff_ / Vgst2Vtm + Vdseff_ * dAbulk_dVg) / Vgst2Vtm;
  dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
  dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff_) / Vgst2Vtm;
  fgche1 = Vgsteff * T0;
  dfgche1_dVg = Vgsteff * dT0_dVg + T0;
  dfgche1_dVd = Vgsteff * dT0_dVd;
  dfgche1_dVb = Vgsteff * dT0_dVb;
  T9 = Vdseff_ / EsatL;
  fgche2 = 1.0 + T9;
  dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
  dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
  dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
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

  if (RELOAD) reload_params(dfb);
  /* output resistance: CLM, DIBL, DITS, SCBE (:9625-9832) */
  if (PP(fprout) <= 0.0) {
    FP = 1.0;
    dFP_dVg = 0.0;
  } else {
    T9 = PP(fprout) * sqrt(Leff) / Vgst2Vtm;
    FP = 1.0 / (1.0 + T9);
    dFP_dVg = FP * FP * T9 / Vgst2Vtm;
  }
  T8 = PP(pvag) / EsatL;
  T9 = T8 * Vgsteff;
  if (T9 > -0.9) {
    PvagTerm = 1.0 + T9;
    dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL);
    dPvagTerm_dVb = -T9 * dEsatL_dVb / EsatL;
    dPvagTerm_dVd = -T9 * dEsatL_dVd / EsatL;
  } else {
    T4 = 1.0 / (17.0 + 20.0 * T9);
    PvagTerm = (0.8 + T9) * T4;
    T4 *= T4;
    dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL) * T4;
    T9 *= T4 / EsatL;
    dPvagTerm_dVb = -T9 * dEsatL_dVb;
    dPvagTerm_dVd = -T9 * dEsatL_dVd;
  }
  if ((PP(pclm) > kMinExp) && (diffVds > 1.0e-10)) {
    T0 = 1.0 + Rds * Idl;
    dT0_dVg = dRds_dVg * Idl + Rds * dIdl_dVg;
    dT0_dVd = Rds * dIdl_dVd;
    dT0_dVb = dRds_dVb * Idl + Rds * dIdl_dVb;
    T2 = Vdsat / Esat;
    T1 = Leff + T2;
    dT1_dVg = (dVdsat_dVg - T2 * dEsatL_dVg / Leff) / Esat;
    dT1_dVd = (dVdsat_dVd - T2 * dEsatL_dVd / Leff) / Esat;
    dT1_dVb = (dVdsat_dVb - T2 * dEsatL_dVb / Leff) / Esat;
    Cclm = FP * PvagTerm * T0 * T1 / (PP(pclm) * PP(litl));
    dCclm_dVg = Cclm * (dFP_dVg / FP + dPvagTerm_dVg / PvagTerm + dT0_dVg / T0 + dT1_dVg / T1);
    dCclm_dVb = Cclm * (dPvagTerm_dVb / PvagTerm + dT0_dVb / T0 + dT1_dVb / T1);
    dCclm_dVd = Cclm * (dPvagTerm_dVd / PvagTerm + dT0_dVd / T0 + dT1_dVd / T1);
    VACLM = Cclm * diffVds;
    dVACLM_dVg = dCclm_dVg * diffVds - dVdseff_dVg * Cclm;
    dVACLM_dVb = dCclm_dVb * diffVds - dVdseff_dVb * Cclm;
    dVACLM_dVd = dCclm_dVd * diffVds + (1.0 - dVdseff_dVd) * Cclm;
  } else {
    VACLM = Cclm = kMaxExp;
    dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
    dCclm_dVd = dCclm_dVg = dCclm_dVb = 0.0;
  }
  if (PP(thetaRout) > kMinExp) {
    T8 = Abulk * Vdsat;
    T0 = Vgst2Vtm * T8;
    dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
    dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
    dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
    T1 = Vgst2Vtm + T8;
    dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
    dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
    dT1_dVd = Abulk * dVdsat_dVd;
    T9 = T1 * T1;
    T2 = PP(thetaRout);
    VADIBL = (Vgst2Vtm - T0 / T1) / T2;
    dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
    dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
    dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
    T7 = PP(pdiblb) * Vbseff;
    if (T7 >= -0.9) {
      T3 = 1.0 / (1.0 + T7);
      VADIBL *= T3;
      dVADIBL_dVg *= T3;
      dVADIBL_dVb = (dVADIBL_dVb - VADIBL * PP(pdiblb)) * T3;
      dVADIBL_dVd *= T3;
    } else {
      T4 = 1.0 / (0.8 + T7);
      T3 = (17.0 + 20.0 * T7) * T4;
      dVADIBL_dVg *= T3;
      dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * PP(pdiblb) * T4 * T4;
      dVADIBL_dVd *= T3;
      VADIBL *= T3;
    }
    dVADIBL_dVg = dVADIBL_dVg * PvagTerm + VADIBL * dPvagTerm_dVg;
    dVADIBL_dVb = dVADIBL_dVb * PvagTerm + VADIBL * dPvagTerm_dVb;
    dVADIBL_dVd = dVADIBL_dVd * PvagTerm + VADIBL * dPvagTerm_dVd;
    VADIBL *= PvagTerm;
  } else {
    VADIBL = kMaxExp;
    dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
  }
  Va = Vasat + VACLM;
  dVa_dVg = dVasat_dVg + dVACLM_dVg;
  dVa_dVb = dVasat_dVb + dVACLM_dVb;
  dVa_dVd = dVasat_dVd + dVACLM_dVd;
  T0 = PP(pditsd) * Vds;
  if (T0 > kExpThreshold) {
    T1 = kMaxExp;
    dT1_dVd = 0;
  } else {
    T1 = exp(T0);
    dT1_dVd = T1 * PP(pditsd);
  }
  if (PP(pdits) > kMinExp) {
    T2 = 1.0 + C.pditsl * Leff;
    VADITS = (1.0 + T2 * T1) / PP(pdits);
    dVADITS_dVg = VADITS * dFP_dVg;
    dVADITS_dVd = FP * T2 * dT1_dVd / PP(pdits);
    VADITS *= FP;
  } else {
    VADITS = kMaxExp;
    dVADITS_dVg = dVADITS_dVd = 0;
  }
  if (PP(pscbe2) > 0.0) {
    if (diffVds > PP(pscbe1) * PP(litl) / kExpThreshold) {
      T0 = PP(pscbe1) * PP(litl) / diffVds;
      VASCBE = Leff * exp(T0) / PP(pscbe2);
      T1 = T0 * VASCBE / diffVds;
      dVASCBE_dVg = T1 * dVdseff_dVg;
      dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
      dVASCBE_dVb = T1 * dVdseff_dVb;
    } else {
      VASCBE = kMaxExp * Leff / PP(pscbe2);
      dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
    }
  } else {
    VASCBE = kMaxExp;
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

  if (RELOAD) reload_params(dfb);
  /* substrate current (:9835-9877) */
  tmp = PP(alpha0) + PP(alpha1) * Leff;
  if ((tmp <= 0.0) || (PP(beta0) <= 0.0)) {
    Isub = Gbd = Gbb = Gbg = 0.0;
  } else {
    T2 = tmp / Leff;
    if (diffVds > PP(beta0) / kExpThreshold) {
      T0 = -PP(beta0) / diffVds;
      T1 = T2 * diffVds * exp(T0);
      T3 = T1 / diffVds * (T0 - 1.0);
      dT1_dVg = T3 * dVdseff_dVg;
      dT1_dVd = T3 * (dVdseff_dVd - 1.0);
      dT1_dVb = T3 * dVdseff_dVb;
    } else {
      T3 = T2 * kMinExp;
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
  csub = Isub;
  gbbs = Gbb;
  gbgs = Gbg;
  gbds = Gbd;

  if (RELOAD) reload_params(dfb);
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
  cdrain = Ids * Vdseff_;

  if (RELOAD) reload_params(dfb);
  /* source-end velocity limit (:9905-9945) */
  if ((C.vtlGiven) && (C.vtl > 0.0)) {
    T12 = 1.0 / Leff / CoxeffWovL;
    T11 = T12 / Vgsteff;
    T10 = -T11 / Vgsteff;
    vs = cdrain * T11;
    dvs_dVg = Gm_ * T11 + cdrain * T10 * dVgsteff_dVg;
    dvs_dVd = Gds * T11 + cdrain * T10 * dVgsteff_dVd;
    dvs_dVb = Gmb * T11 + cdrain * T10 * dVgsteff_dVb;
    T0 = 2 * kMM;
    T1 = vs / (PP(vtl) * PP(tfactor));
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
    Gm_ += cdrain * dFsevl_dVg;
    Gmb *= Fsevl;
    Gmb += cdrain * dFsevl_dVb;
    Gds *= Fsevl;
    Gds += cdrain * dFsevl_dVd;
    cdrain *= Fsevl;
  }
  gds = Gds;
  gm = Gm_;
  gmbs = Gmb;
  IdovVds = Ids;
  if (IdovVds <= 1.0e-9) IdovVds = 1.0e-9;
  if (RELOAD) reload_params(dfb);
  /* gate resistance conductances (:9953-9988); without rgateMod > 1 or NQS
     they stay at the zeros setupFVectorVars left */
  if (X && ((irgateMod > 1) || (itrnqsMod != 0) || (iacnqsMod != 0))) {
    T9 = XP(xrcrg2) * C.vtm;
    T0 = T9 * beta;
    dT0_dVd = (dbeta_dVd + dbeta_dVg * dVgsteff_dVd) * T9;
    dT0_dVb = (dbeta_dVb + dbeta_dVg * dVgsteff_dVb) * T9;
    dT0_dVg = dbeta_dVg * T9;
    gcrg = XP(xrcrg1) * (T0 + Ids);
    gcrgd = XP(xrcrg1) * (dT0_dVd + tmp1);
    gcrgb = XP(xrcrg1) * (dT0_dVb + tmp2) * dVbseff_dVb;
    gcrgg = XP(xrcrg1) * (dT0_dVg + tmp3) * dVgsteff_dVg;
    if (nf != 1.0) {
      gcrg *= nf;
      gcrgg *= nf;
      gcrgd *= nf;
      gcrgb *= nf;
    }
    if (irgateMod == 2) {
      T10 = XP(grgeltd) * XP(grgeltd);
      T11 = XP(grgeltd) + gcrg;
      gcrg = XP(grgeltd) * gcrg / T11;
      T12 = T10 / T11 / T11;
      gcrgg *= T12;
      gcrgd *= T12;
      gcrgb *= T12;
    }
    gcrgs = -(gcrgg + gcrgd + gcrgb);
  } else {
    gcrg = gcrgd = gcrgg = gcrgs = gcrgb = 0.0;
  }
  if (RELOAD) reload_params(dfb);
  /* bias-dependent external source/drain resistance (:9991-10069) */
  if (X && rdsMod) {
    double dT0_dvg, dT1_dvb, dT3_dvg, dT3_dvb, Rs, dRs_dvg, dRs_dvb, Rd, dRd_dvg, dRd_dvb;
    double dgstot_dvd, dgstot_dvg, dgstot_dvb, dgstot_dvs, dgdtot_dvd, dgdtot_dvg, dgdtot_dvb, dgdtot_dvs;
    T0 = vgs - XP(vfbsd);
    T1 = sqrt(T0 * T0 + 1.0e-4);
    vgs_eff = 0.5 * (T0 + T1);
    dvgs_eff_dvg = vgs_eff / T1;
    T0 = 1.0 + PP(prwg) * vgs_eff;
    dT0_dvg = -PP(prwg) / T0 / T0 * dvgs_eff_dvg;
    T1 = -PP(prwb) * vbs;
    dT1_dvb = -PP(prwb);
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    dT3_dvg = T3 / (T3 - T2);
    dT3_dvb = dT3_dvg * dT1_dvb;
    dT3_dvg *= dT0_dvg;
    T4 = XP(rs0) * 0.5;
    Rs = XP(rswmin) + T3 * T4;
    dRs_dvg = T4 * dT3_dvg;
    dRs_dvb = T4 * dT3_dvb;
    T0 = 1.0 + IV(sourceConductance) * Rs;
    gstot = IV(sourceConductance) / T0;
    T0 = -gstot * gstot;
    dgstot_dvd = 0.0;
    dgstot_dvg = T0 * dRs_dvg;
    dgstot_dvb = T0 * dRs_dvb;
    dgstot_dvs = -(dgstot_dvg + dgstot_dvb + dgstot_dvd);
    T0 = vgd - XP(vfbsd);
    T1 = sqrt(T0 * T0 + 1.0e-4);
    vgd_eff = 0.5 * (T0 + T1);
    dvgd_eff_dvg = vgd_eff / T1;
    T0 = 1.0 + PP(prwg) * vgd_eff;
    dT0_dvg = -PP(prwg) / T0 / T0 * dvgd_eff_dvg;
    T1 = -PP(prwb) * vbd;
    dT1_dvb = -PP(prwb);
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    dT3_dvg = T3 / (T3 - T2);
    dT3_dvb = dT3_dvg * dT1_dvb;
    dT3_dvg *= dT0_dvg;
    T4 = XP(rd0) * 0.5;
    Rd = XP(rdwmin) + T3 * T4;
    dRd_dvg = T4 * dT3_dvg;
    dRd_dvb = T4 * dT3_dvb;
    T0 = 1.0 + IV(drainConductance) * Rd;
    gdtot = IV(drainConductance) / T0;
    T0 = -gdtot * gdtot;
    dgdtot_dvs = 0.0;
    dgdtot_dvg = T0 * dRd_dvg;
    dgdtot_dvb = T0 * dRd_dvb;
    dgdtot_dvd = -(dgdtot_dvg + dgdtot_dvb + dgdtot_dvs);
    gstotd = vses * dgstot_dvd;
    gstotg = vses * dgstot_dvg;
    gstots = vses * dgstot_dvs;
    gstotb = vses * dgstot_dvb;
    T2 = vdes - vds;
    gdtotd = T2 * dgdtot_dvd;
    gdtotg = T2 * dgdtot_dvg;
    gdtots = T2 * dgdtot_dvs;
    gdtotb = T2 * dgdtot_dvb;
  } else {
    gstot = gstotd = gstotg = 0.0;
    gstots = gstotb = 0.0;
    gdtot = gdtotd = gdtotg = 0.0;
    gdtots = gdtotb = 0.0;
  }


  if (RELOAD) reload_params(dfb);
  /* GIDL (:10074-10126) */
  if (X && mtrlMod) {
    T0 = XP(epsrsub) * toxe / epsrox;
    T1 = (vds - vgs_eff - PP(egidl) + XP(vfbsd)) / T0;
  } else {
    T0 = 3.0 * toxe;
    T1 = (vds - vgs_eff - PP(egidl)) / T0;
  }
  if ((PP(agidl) <= 0.0) || (PP(bgidl) <= 0.0) || (T1 <= 0.0) || (PP(cgidl) <= 0.0) || (vbd > 0.0)) {
    Igidl = Ggidld = Ggidlg = Ggidlb = 0.0;
  } else {
    dT1_dVd = 1.0 / T0;
    dT1_dVg = -dvgs_eff_dvg * dT1_dVd;
    T2 = PP(bgidl) / T1;
    if (T2 < 100.0) {
      Igidl = PP(agidl) * PP(weffCJ) * T1 * exp(-T2);
      T3 = Igidl * (1.0 + T2) / T1;
      Ggidld = T3 * dT1_dVd;
      Ggidlg = T3 * dT1_dVg;
    } else {
      Igidl = PP(agidl) * PP(weffCJ) * 3.720075976e-44;
      Ggidld = Igidl * dT1_dVd;
      Ggidlg = Igidl * dT1_dVg;
      Igidl *= T1;
    }
    T4 = vbd * vbd;
    T5 = -vbd * T4;
    T6 = PP(cgidl) + T5;
    T7 = T5 / T6;
    T8 = 3.0 * PP(cgidl) * T4 / T6 / T6;
    Ggidld = Ggidld * T7 + Igidl * T8;
    Ggidlg = Ggidlg * T7;
    Ggidlb = -Igidl * T8;
    Igidl *= T7;
  }
  ggidld = Ggidld;
  ggidlg = Ggidlg;
  ggidlb = Ggidlb;
  if (RELOAD) reload_params(dfb);
  /* GISL (:10132-10179) */
  if (X && mtrlMod) T1 = (-vds - vgd_eff - PP(egisl) + XP(vfbsd)) / T0;
  else T1 = (-vds - vgd_eff - PP(egisl)) / T0;
  if ((PP(agisl) <= 0.0) || (PP(bgisl) <= 0.0) || (T1 <= 0.0) || (PP(cgisl) <= 0.0) || (vbs > 0.0)) {
    Igisl = Ggisls = Ggislg = Ggislb = 0.0;
  } else {
    dT1_dVd = 1.0 / T0;
    dT1_dVg = -dvgd_eff_dvg * dT1_dVd;
    T2 = PP(bgisl) / T1;
    if (T2 < 100.0) {
      Igisl = PP(agisl) * PP(weffCJ) * T1 * exp(-T2);
      T3 = Igisl * (1.0 + T2) / T1;
      Ggisls = T3 * dT1_dVd;
      Ggislg = T3 * dT1_dVg;
    } else {
      Igisl = PP(agisl) * PP(weffCJ) * 3.720075976e-44;
      Ggisls = Igisl * dT1_dVd;
      Ggislg = Igisl * dT1_dVg;
      Igisl *= T1;
    }
    T4 = vbs * vbs;
    T5 = -vbs * T4;
    T6 = PP(cgisl) + T5;
    T7 = T5 / T6;
    T8 = 3.0 * PP(cgisl) * T4 / T6 / T6;
    Ggisls = Ggisls * T7 + Igisl * T8;
    Ggislg = Ggislg * T7;
    Ggislb = -Igisl * T8;
    Igisl *= T7;
  }
  ggisls = Ggisls;
  ggislg = Ggislg;
  ggislb = Ggislb;
  if (RELOAD) reload_params(dfb);
  if constexpr (X) {
    double Vfb, Voxacc, dVoxacc_dVg, dVoxacc_dVb, Voxdepinv, dVoxdepinv_dVg, dVoxdepinv_dVd;
    double dVoxdepinv_dVb, VxNVt, ExpVxNVt, Vaux, dVaux_dVg, dVaux_dVd, dVaux_dVb;
    double dT7_dVg, dT7_dVd, dT7_dVb, Igc, dIgc_dVg, dIgc_dVd, dIgc_dVb, Pigcd, dPigcd_dVg;
    double dPigcd_dVd, dPigcd_dVb, dIgcs_dVg, dIgcs_dVd, dIgcs_dVb, dIgcd_dVg, dIgcd_dVd, dIgcd_dVb;
    double dIgs_dVg, dIgs_dVs, dIgd_dVg, dIgd_dVd, Igbacc, dIgbacc_dVg, dIgbacc_dVb;
    double Igbinv, dIgbinv_dVg, dIgbinv_dVd, dIgbinv_dVb;
    /* gate tunnelling current (:10182-10621) */
    Vfb = 0.0;
    Voxacc = dVoxacc_dVg = dVoxacc_dVb = 0.0;
    Voxdepinv = dVoxdepinv_dVg = dVoxdepinv_dVd = dVoxdepinv_dVb = 0.0;
    if ((igcMod != 0) || (igbMod != 0)) {
      Vfb = IV(vfbzb);
      V3 = Vfb - Vgs_eff + Vbseff - kDelta3;
      if (Vfb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * kDelta3 * Vfb);
      else T0 = sqrt(V3 * V3 + 4.0 * kDelta3 * Vfb);
      T1 = 0.5 * (1.0 + V3 / T0);
      Vfbeff = Vfb - 0.5 * (V3 + T0);
      dVfbeff_dVg = T1 * dVgs_eff_dVg;
      dVfbeff_dVb = -T1;
      Voxacc = Vfb - Vfbeff;
      dVoxacc_dVg = -dVfbeff_dVg;
      dVoxacc_dVb = -dVfbeff_dVb;
      if (Voxacc < 0.0) Voxacc = dVoxacc_dVg = dVoxacc_dVb = 0.0;
      T0 = 0.5 * PP(k1ox);
      T3 = Vgs_eff - Vfbeff - Vbseff - Vgsteff;
      if (PP(k1ox) == 0.0) {
        Voxdepinv = dVoxdepinv_dVg = dVoxdepinv_dVd = dVoxdepinv_dVb = 0.0;
      } else if (T3 < 0.0) {
        Voxdepinv = -T3;
        dVoxdepinv_dVg = -dVgs_eff_dVg + dVfbeff_dVg + dVgsteff_dVg;
        dVoxdepinv_dVd = dVgsteff_dVd;
        dVoxdepinv_dVb = dVfbeff_dVb + 1.0 + dVgsteff_dVb;
      } else {
        T1 = sqrt(T0 * T0 + T3);
        T2 = T0 / T1;
        Voxdepinv = PP(k1ox) * (T1 - T0);
        dVoxdepinv_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg);
        dVoxdepinv_dVd = -T2 * dVgsteff_dVd;
        dVoxdepinv_dVb = -T2 * (dVfbeff_dVb + 1.0 + dVgsteff_dVb);
      }
      Voxdepinv += Vgsteff;
      dVoxdepinv_dVg += dVgsteff_dVg;
      dVoxdepinv_dVd += dVgsteff_dVd;
      dVoxdepinv_dVb += dVgsteff_dVb;
    }
    if (C.tempMod < 2) tmp = Vtm;
    else tmp = Vtm0;
    if (igcMod) {
      T0 = tmp * XP(nigc);
      VxNVt = 0.0;
      Vaux = dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
      if (igcMod == 1) {
        VxNVt = (Vgs_eff - C.dtype * IV(vth0)) / T0;
        if (VxNVt > kExpThreshold) {
          Vaux = Vgs_eff - C.dtype * IV(vth0);
          dVaux_dVg = dVgs_eff_dVg;
          dVaux_dVd = 0.0;
          dVaux_dVb = 0.0;
        }
      } else if (igcMod == 2) {
        VxNVt = (Vgs_eff - von) / T0;
        if (VxNVt > kExpThreshold) {
          Vaux = Vgs_eff - von;
          dVaux_dVg = dVgs_eff_dVg;
          dVaux_dVd = -dVth_dVd;
          dVaux_dVb = -dVth_dVb;
        }
      }
      if (VxNVt < -kExpThreshold) {
        Vaux = T0 * log(1.0 + kMinExp);
        dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
      } else if ((VxNVt >= -kExpThreshold) && (VxNVt <= kExpThreshold)) {
        ExpVxNVt = exp(VxNVt);
        Vaux = T0 * log(1.0 + ExpVxNVt);
        dVaux_dVg = ExpVxNVt / (1.0 + ExpVxNVt);
        if (igcMod == 1) {
          dVaux_dVd = 0.0;
          dVaux_dVb = 0.0;
        } else if (igcMod == 2) {
          dVaux_dVd = -dVgs_eff_dVg * dVth_dVd;
          dVaux_dVb = -dVgs_eff_dVg * dVth_dVb;
        }
        dVaux_dVg *= dVgs_eff_dVg;
      }
      T2 = Vgs_eff * Vaux;
      dT2_dVg = dVgs_eff_dVg * Vaux + Vgs_eff * dVaux_dVg;
      dT2_dVd = Vgs_eff * dVaux_dVd;
      dT2_dVb = Vgs_eff * dVaux_dVb;
      T11 = XP(Aechvb);
      T12 = XP(Bechvb);
      T3 = XP(aigc) * XP(cigc) - XP(bigc);
      T4 = XP(bigc) * XP(cigc);
      T5 = T12 * (XP(aigc) + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
      if (T5 > kExpThreshold) {
        T6 = kMaxExp;
        dT6_dVg = dT6_dVd = dT6_dVb = 0.0;
      } else if (T5 < -kExpThreshold) {
        T6 = kMinExp;
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
      if (((xf >> XB_PIGCD) & 1)) {
        Pigcd = XP(pigcd);
        dPigcd_dVg = dPigcd_dVd = dPigcd_dVb = 0.0;
      } else {
        T11 = XP(Bechvb) * toxe;
        T12 = Vgsteff + 1.0e-20;
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
      if (T7 > kExpThreshold) {
        T9 = kMaxExp;
        dT9_dVg = dT9_dVd = dT9_dVb = 0.0;
      } else if (T7 < -kExpThreshold) {
        T9 = kMinExp;
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
      Igcs = Igc * T10;
      dIgcs_dVg = dIgc_dVg * T10 + Igc * dT10_dVg;
      dIgcs_dVd = dIgc_dVd * T10 + Igc * dT10_dVd;
      dIgcs_dVb = dIgc_dVb * T10 + Igc * dT10_dVb;
      T1 = T9 - 1.0 - 1.0e-4;
      T10 = (T7 * T9 - T1) / T8;
      dT10_dVg = (dT7_dVg * T9 + (T7 - 1.0) * dT9_dVg - T10 * dT8_dVg) / T8;
      dT10_dVd = (dT7_dVd * T9 + (T7 - 1.0) * dT9_dVd - T10 * dT8_dVd) / T8;
      dT10_dVb = (dT7_dVb * T9 + (T7 - 1.0) * dT9_dVb - T10 * dT8_dVb) / T8;
      Igcd = Igc * T10;
      dIgcd_dVg = dIgc_dVg * T10 + Igc * dT10_dVg;
      dIgcd_dVd = dIgc_dVd * T10 + Igc * dT10_dVd;
      dIgcd_dVb = dIgc_dVb * T10 + Igc * dT10_dVb;
      gIgcsg = dIgcs_dVg;
      gIgcsd = dIgcs_dVd;
      gIgcsb = dIgcs_dVb * dVbseff_dVb;
      gIgcdg = dIgcd_dVg;
      gIgcdd = dIgcd_dVd;
      gIgcdb = dIgcd_dVb * dVbseff_dVb;
      /* source/drain edge tunnelling: rewrites the vgs_eff / vgd_eff members */
      T0 = vgs - (XP(vfbsd) + XP(vfbsdoff));
      vgs_eff = sqrt(T0 * T0 + 1.0e-4);
      dvgs_eff_dvg = T0 / vgs_eff;
      T2 = vgs * vgs_eff;
      dT2_dVg = vgs * dvgs_eff_dvg + vgs_eff;
      T11 = XP(AechvbEdgeS);
      T12 = XP(BechvbEdge);
      T3 = XP(aigs) * XP(cigs) - XP(bigs);
      T4 = XP(bigs) * XP(cigs);
      T5 = T12 * (XP(aigs) + T3 * vgs_eff - T4 * vgs_eff * vgs_eff);
      if (T5 > kExpThreshold) {
        T6 = kMaxExp;
        dT6_dVg = 0.0;
      } else if (T5 < -kExpThreshold) {
        T6 = kMinExp;
        dT6_dVg = 0.0;
      } else {
        T6 = exp(T5);
        dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * vgs_eff) * dvgs_eff_dvg;
      }
      Igs = T11 * T2 * T6;
      dIgs_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
      dIgs_dVs = -dIgs_dVg;
      T0 = vgd - (XP(vfbsd) + XP(vfbsdoff));
      vgd_eff = sqrt(T0 * T0 + 1.0e-4);
      dvgd_eff_dvg = T0 / vgd_eff;
      T2 = vgd * vgd_eff;
      dT2_dVg = vgd * dvgd_eff_dvg + vgd_eff;
      T11 = XP(AechvbEdgeD);
      T3 = XP(aigd) * XP(cigd) - XP(bigd);
      T4 = XP(bigd) * XP(cigd);
      T5 = T12 * (XP(aigd) + T3 * vgd_eff - T4 * vgd_eff * vgd_eff);
      if (T5 > kExpThreshold) {
        T6 = kMaxExp;
        dT6_dVg = 0.0;
      } else if (T5 < -kExpThreshold) {
        T6 = kMinExp;
        dT6_dVg = 0.0;
      } else {
        T6 = exp(T5);
        dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * vgd_eff) * dvgd_eff_dvg;
      }
      Igd = T11 * T2 * T6;
      dIgd_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
      dIgd_dVd = -dIgd_dVg;
      gIgsg = dIgs_dVg;
      gIgss = dIgs_dVs;
      gIgdg = dIgd_dVg;
      gIgdd = dIgd_dVd;
    } else {
      Igcs = gIgcsg = gIgcsd = gIgcsb = 0.0;
      Igcd = gIgcdg = gIgcdd = gIgcdb = 0.0;
      Igs = gIgsg = gIgss = 0.0;
      Igd = gIgdg = gIgdd = 0.0;
    }
    if (igbMod) {
      T0 = tmp * XP(nigbacc);
      T1 = -Vgs_eff + Vbseff + Vfb;
      VxNVt = T1 / T0;
      if (VxNVt > kExpThreshold) {
        Vaux = T1;
        dVaux_dVg = -dVgs_eff_dVg;
        dVaux_dVb = 1.0;
      } else if (VxNVt < -kExpThreshold) {
        Vaux = T0 * log(1.0 + kMinExp);
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
      T11 = 4.97232e-7 * PP(weff) * PP(leff) * XP(ToxRatio);
      T12 = -7.45669e11 * toxe;
      T3 = XP(aigbacc) * XP(cigbacc) - XP(bigbacc);
      T4 = XP(bigbacc) * XP(cigbacc);
      T5 = T12 * (XP(aigbacc) + T3 * Voxacc - T4 * Voxacc * Voxacc);
      if (T5 > kExpThreshold) {
        T6 = kMaxExp;
        dT6_dVg = dT6_dVb = 0.0;
      } else if (T5 < -kExpThreshold) {
        T6 = kMinExp;
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
      T0 = tmp * XP(nigbinv);
      T1 = Voxdepinv - XP(eigbinv);
      VxNVt = T1 / T0;
      if (VxNVt > kExpThreshold) {
        Vaux = T1;
        dVaux_dVg = dVoxdepinv_dVg;
        dVaux_dVd = dVoxdepinv_dVd;
        dVaux_dVb = dVoxdepinv_dVb;
      } else if (VxNVt < -kExpThreshold) {
        Vaux = T0 * log(1.0 + kMinExp);
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
      T3 = XP(aigbinv) * XP(cigbinv) - XP(bigbinv);
      T4 = XP(bigbinv) * XP(cigbinv);
      T5 = T12 * (XP(aigbinv) + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
      if (T5 > kExpThreshold) {
        T6 = kMaxExp;
        dT6_dVg = dT6_dVd = dT6_dVb = 0.0;
      } else if (T5 < -kExpThreshold) {
        T6 = kMinExp;
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
      Igb = Igbinv + Igbacc;
      gIgbg = dIgbinv_dVg + dIgbacc_dVg;
      gIgbd = dIgbinv_dVd;
      gIgbb = (dIgbinv_dVb + dIgbacc_dVb) * dVbseff_dVb;
    } else {
      Igb = gIgbg = gIgbd = gIgbs = gIgbb = 0.0;
    }
  } else {
    /* igcMod = igbMod = 0 (:10482-10488, :10617-10620) */
    Igcs = gIgcsg = gIgcsd = gIgcsb = 0.0;
    Igcd = gIgcdg = gIgcdd = gIgcdb = 0.0;
    Igs = gIgsg = gIgss = 0.0;
    Igd = gIgdg = gIgdd = 0.0;
    Igb = gIgbg = gIgbd = gIgbs = gIgbb = 0.0;
  }
  if (nf != 1.0) {
    cdrain *= nf;
    gds *= nf;
    gm *= nf;
    gmbs *= nf;
    IdovVds *= nf;
    gbbs *= nf;
    gbgs *= nf;
    gbds *= nf;
    csub *= nf;
    Igidl *= nf;
    ggidld *= nf;
    ggidlg *= nf;
    ggidlb *= nf;
    Igisl *= nf;
    ggisls *= nf;
    ggislg *= nf;
    ggislb *= nf;
    Igcs *= nf;
    gIgcsg *= nf;
    gIgcsd *= nf;
    gIgcsb *= nf;
    Igcd *= nf;
    gIgcdg *= nf;
    gIgcdd *= nf;
    gIgcdb *= nf;
    Igs *= nf;
    gIgsg *= nf;
    gIgss *= nf;
    Igd *= nf;
    gIgdg *= nf;
    gIgdd *= nf;
    Igb *= nf;
    gIgbg *= nf;
    gIgbd *= nf;
    gIgbb *= nf;
  }
  ggidls = -(ggidld + ggidlg + ggidlb);
  ggisld = -(ggisls + ggislg + ggislb);
  gIgbs = -(gIgbg + gIgbd + gIgbb);
  gIgcss = -(gIgcsg + gIgcsd + gIgcsb);
  gIgcds = -(gIgcdg + gIgcdd + gIgcdb);
  cd = cdrain;

  if (RELOAD) reload_params(dfb);
  /* noise charge (:10674-10703): rewrites the Abulk and Vdseff members */
  if (C.tnoiMod == 0) {
    Abulk = Abulk0 * PP(abulkCVfactor);
    Vdsat = Vgsteff / Abulk;
    T0 = Vdsat - Vds - kDelta4;
    T1 = sqrt(T0 * T0 + 4.0 * kDelta4 * Vdsat);
    if (T0 >= 0.0) {
      Vdseff = Vdsat - 0.5 * (T0 + T1);
    } else {
      T3 = (kDelta4 + kDelta4) / (T1 - T0);
      T4 = 1.0 - T3;
      T5 = Vdsat * T3 / (T1 - T0);
      Vdseff = Vdsat * T4;
    }
    if (Vds == 0.0) Vdseff = 0.0;
    T0 = Abulk * Vdseff;
    T1 = 12.0 * (Vgsteff - 0.5 * T0 + 1.0e-20);
    T2 = Vdseff / T1;
    T3 = T0 * T2;
    qinv = Coxeff * PP(weffCV) * nf * PP(leffCV) * (Vgsteff - 0.5 * T0 + Abulk * T3);
  }

  if (RELOAD) reload_params(dfb);
  /* intrinsic charges (:10707-11611) */
  if ((C.xpart < 0) || (!ChargeComputationNeeded)) {
    qgate = qdrn = qsrc = qbulk = 0.0;
    cggb = cgsb = cgdb = 0.0;
    cdgb = cdsb = cddb = 0.0;
    cbgb = cbsb = cbdb = 0.0;
    csgb = cssb = csdb = 0.0;
    cgbb = csbb = cdbb = cbbb = 0.0;
    cqdb = cqsb = cqgb = cqbb = 0.0;
    gtau = 0.0;
  } else if (X && capMod == 0) {
    double Arg1, One_Third_CoxWL, Two_Third_CoxWL, Alphaz, dAlphaz_dVg, dAlphaz_dVb;
    double Vfb, dVgst_dVb, dVgst_dVg;
    /* capMod = 0 (:10718-11012) */
    if (Vbseff < 0.0) {
      Vbseff = Vbs;
      dVbseff_dVb = 1.0;
    } else {
      Vbseff = PP(phi) - Phis;
      dVbseff_dVb = -dPhis_dVb;
    }
    Vfb = XP(vfbcv);
    Vth = Vfb + PP(phi) + PP(k1ox) * sqrtPhis;
    Vgst = Vgs_eff - Vth;
    dVth_dVb = PP(k1ox) * dsqrtPhis_dVb;
    dVgst_dVb = -dVth_dVb;
    dVgst_dVg = dVgs_eff_dVg;
    CoxWL = C.coxe * PP(weffCV) * PP(leffCV) * nf;
    Arg1 = Vgs_eff - Vbseff - Vfb;
    if (Arg1 <= 0.0) {
      qgate = CoxWL * Arg1;
      qbulk = -qgate;
      qdrn = 0.0;
      cggb = CoxWL * dVgs_eff_dVg;
      cgdb = 0.0;
      cgsb = CoxWL * (dVbseff_dVb - dVgs_eff_dVg);
      cdgb = 0.0;
      cddb = 0.0;
      cdsb = 0.0;
      cbgb = -CoxWL * dVgs_eff_dVg;
      cbdb = 0.0;
      cbsb = -cgsb;
    } else if (Vgst <= 0.0) {
      T1 = 0.5 * PP(k1ox);
      T2 = sqrt(T1 * T1 + Arg1);
      qgate = CoxWL * PP(k1ox) * (T2 - T1);
      qbulk = -qgate;
      qdrn = 0.0;
      T0 = CoxWL * T1 / T2;
      cggb = T0 * dVgs_eff_dVg;
      cgdb = 0.0;
      cgsb = T0 * (dVbseff_dVb - dVgs_eff_dVg);
      cdgb = 0.0;
      cddb = 0.0;
      cdsb = 0.0;
      cbgb = -cggb;
      cbdb = 0.0;
      cbsb = -cgsb;
    } else {
      One_Third_CoxWL = CoxWL / 3.0;
      Two_Third_CoxWL = 2.0 * One_Third_CoxWL;
      AbulkCV = Abulk0 * PP(abulkCVfactor);
      dAbulkCV_dVb = PP(abulkCVfactor) * dAbulk0_dVb;
      Vdsat = Vgst / AbulkCV;
      dVdsat_dVg = dVgs_eff_dVg / AbulkCV;
      dVdsat_dVb = -(Vdsat * dAbulkCV_dVb + dVth_dVb) / AbulkCV;
      if (C.xpart > 0.5) {
        /* 0/100 partition */
        if (Vdsat <= Vds) {
          T1 = Vdsat / 3.0;
          qgate = CoxWL * (Vgs_eff - Vfb - PP(phi) - T1);
          T2 = -Two_Third_CoxWL * Vgst;
          qbulk = -(qgate + T2);
          qdrn = 0.0;
          cggb = One_Third_CoxWL * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
          T2 = -One_Third_CoxWL * dVdsat_dVb;
          cgsb = -(cggb + T2);
          cgdb = 0.0;
          cdgb = 0.0;
          cddb = 0.0;
          cdsb = 0.0;
          cbgb = -(cggb - Two_Third_CoxWL * dVgs_eff_dVg);
          T3 = -(T2 + Two_Third_CoxWL * dVth_dVb);
          cbsb = -(cbgb + T3);
          cbdb = 0.0;
        } else {
          Alphaz = Vgst / Vdsat;
          T1 = 2.0 * Vdsat - Vds;
          T2 = Vds / (3.0 * T1);
          T3 = T2 * Vds;
          T9 = 0.25 * CoxWL;
          T4 = T9 * Alphaz;
          T7 = 2.0 * Vds - T1 - 3.0 * T3;
          T8 = T3 - T1 - 2.0 * Vds;
          qgate = CoxWL * (Vgs_eff - Vfb - PP(phi) - 0.5 * (Vds - T3));
          T10 = T4 * T8;
          qdrn = T4 * T7;
          qbulk = -(qgate + qdrn + T10);
          T5 = T3 / T1;
          cggb = CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
          T11 = -CoxWL * T5 * dVdsat_dVb;
          cgdb = CoxWL * (T2 - 0.5 + 0.5 * T5);
          cgsb = -(cggb + T11 + cgdb);
          T6 = 1.0 / Vdsat;
          dAlphaz_dVg = T6 * (1.0 - Alphaz * dVdsat_dVg);
          dAlphaz_dVb = -T6 * (dVth_dVb + Alphaz * dVdsat_dVb);
          T7 = T9 * T7;
          T8 = T9 * T8;
          T9 = 2.0 * T4 * (1.0 - 3.0 * T5);
          cdgb = (T7 * dAlphaz_dVg - T9 * dVdsat_dVg) * dVgs_eff_dVg;
          T12 = T7 * dAlphaz_dVb - T9 * dVdsat_dVb;
          cddb = T4 * (3.0 - 6.0 * T2 - 3.0 * T5);
          cdsb = -(cdgb + T12 + cddb);
          T9 = 2.0 * T4 * (1.0 + T5);
          T10 = (T8 * dAlphaz_dVg - T9 * dVdsat_dVg) * dVgs_eff_dVg;
          T11 = T8 * dAlphaz_dVb - T9 * dVdsat_dVb;
          T12 = T4 * (2.0 * T2 + T5 - 1.0);
          T0 = -(T10 + T11 + T12);
          cbgb = -(cggb + cdgb + T10);
          cbdb = -(cgdb + cddb + T12);
          cbsb = -(cgsb + cdsb + T0);
        }
      } else if (C.xpart < 0.5) {
        /* 40/60 partition */
        if (Vds >= Vdsat) {
          T1 = Vdsat / 3.0;
          qgate = CoxWL * (Vgs_eff - Vfb - PP(phi) - T1);
          T2 = -Two_Third_CoxWL * Vgst;
          qbulk = -(qgate + T2);
          qdrn = 0.4 * T2;
          cggb = One_Third_CoxWL * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
          T2 = -One_Third_CoxWL * dVdsat_dVb;
          cgsb = -(cggb + T2);
          cgdb = 0.0;
          T3 = 0.4 * Two_Third_CoxWL;
          cdgb = -T3 * dVgs_eff_dVg;
          cddb = 0.0;
          T4 = T3 * dVth_dVb;
          cdsb = -(T4 + cdgb);
          cbgb = -(cggb - Two_Third_CoxWL * dVgs_eff_dVg);
          T3 = -(T2 + Two_Third_CoxWL * dVth_dVb);
          cbsb = -(cbgb + T3);
          cbdb = 0.0;
        } else {
          Alphaz = Vgst / Vdsat;
          T1 = 2.0 * Vdsat - Vds;
          T2 = Vds / (3.0 * T1);
          T3 = T2 * Vds;
          T9 = 0.25 * CoxWL;
          T4 = T9 * Alphaz;
          qgate = CoxWL * (Vgs_eff - Vfb - PP(phi) - 0.5 * (Vds - T3));
          T5 = T3 / T1;
          cggb = CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
          tmp = -CoxWL * T5 * dVdsat_dVb;
          cgdb = CoxWL * (T2 - 0.5 + 0.5 * T5);
          cgsb = -(cggb + cgdb + tmp);
          T6 = 1.0 / Vdsat;
          dAlphaz_dVg = T6 * (1.0 - Alphaz * dVdsat_dVg);
          dAlphaz_dVb = -T6 * (dVth_dVb + Alphaz * dVdsat_dVb);
          T6 = 8.0 * Vdsat * Vdsat - 6.0 * Vdsat * Vds + 1.2 * Vds * Vds;
          T8 = T2 / T1;
          T7 = Vds - T1 - T8 * T6;
          qdrn = T4 * T7;
          T7 *= T9;
          tmp = T8 / T1;
          tmp1 = T4 * (2.0 - 4.0 * tmp * T6 + T8 * (16.0 * Vdsat - 6.0 * Vds));
          cdgb = (T7 * dAlphaz_dVg - tmp1 * dVdsat_dVg) * dVgs_eff_dVg;
          T10 = T7 * dAlphaz_dVb - tmp1 * dVdsat_dVb;
          cddb = T4 * (2.0 - (1.0 / (3.0 * T1 * T1) + 2.0 * tmp) * T6
                          + T8 * (6.0 * Vdsat - 2.4 * Vds));
          cdsb = -(cdgb + T10 + cddb);
          T7 = 2.0 * (T1 + T3);
          qbulk = -(qgate - T4 * T7);
          T7 *= T9;
          T0 = 4.0 * T4 * (1.0 - T5);
          T12 = (-T7 * dAlphaz_dVg - cdgb - T0 * dVdsat_dVg) * dVgs_eff_dVg;
          T11 = -T7 * dAlphaz_dVb - T10 - T0 * dVdsat_dVb;
          T10 = -4.0 * T4 * (T2 - 0.5 + 0.5 * T5) - cddb;
          tmp = -(T10 + T11 + T12);
          cbgb = -(cggb + cdgb + T12);
          cbdb = -(cgdb + cddb + T10);
          cbsb = -(cgsb + cdsb + tmp);
        }
      } else {
        /* 50/50 partition */
        if (Vds >= Vdsat) {
          T1 = Vdsat / 3.0;
          qgate = CoxWL * (Vgs_eff - Vfb - PP(phi) - T1);
          T2 = -Two_Third_CoxWL * Vgst;
          qbulk = -(qgate + T2);
          qdrn = 0.5 * T2;
          cggb = One_Third_CoxWL * (3.0 - dVdsat_dVg) * dVgs_eff_dVg;
          T2 = -One_Third_CoxWL * dVdsat_dVb;
          cgsb = -(cggb + T2);
          cgdb = 0.0;
          cdgb = -One_Third_CoxWL * dVgs_eff_dVg;
          cddb = 0.0;
          T4 = One_Third_CoxWL * dVth_dVb;
          cdsb = -(T4 + cdgb);
          cbgb = -(cggb - Two_Third_CoxWL * dVgs_eff_dVg);
          T3 = -(T2 + Two_Third_CoxWL * dVth_dVb);
          cbsb = -(cbgb + T3);
          cbdb = 0.0;
        } else {
          Alphaz = Vgst / Vdsat;
          T1 = 2.0 * Vdsat - Vds;
          T2 = Vds / (3.0 * T1);
          T3 = T2 * Vds;
          T9 = 0.25 * CoxWL;
          T4 = T9 * Alphaz;
          qgate = CoxWL * (Vgs_eff - Vfb - PP(phi) - 0.5 * (Vds - T3));
          T5 = T3 / T1;
          cggb = CoxWL * (1.0 - T5 * dVdsat_dVg) * dVgs_eff_dVg;
          tmp = -CoxWL * T5 * dVdsat_dVb;
          cgdb = CoxWL * (T2 - 0.5 + 0.5 * T5);
          cgsb = -(cggb + cgdb + tmp);
          T6 = 1.0 / Vdsat;
          dAlphaz_dVg = T6 * (1.0 - Alphaz * dVdsat_dVg);
          dAlphaz_dVb = -T6 * (dVth_dVb + Alphaz * dVdsat_dVb);
          T7 = T1 + T3;
          qdrn = -T4 * T7;
          qbulk = -(qgate + qdrn + qdrn);
          T7 *= T9;
          T0 = T4 * (2.0 * T5 - 2.0);
          cdgb = (T0 * dVdsat_dVg - T7 * dAlphaz_dVg) * dVgs_eff_dVg;
          T12 = T0 * dVdsat_dVb - T7 * dAlphaz_dVb;
          cddb = T4 * (1.0 - 2.0 * T2 - T5);
          cdsb = -(cdgb + T12 + cddb);
          cbgb = -(cggb + 2.0 * cdgb);
          cbdb = -(cgdb + 2.0 * cddb);
          cbsb = -(cgsb + 2.0 * cdsb);
        }
      }
    }
  } else {
    if (Vbseff < 0.0) {
      VbseffCV = Vbseff;
      dVbseffCV_dVb = 1.0;
    } else {
      VbseffCV = PP(phi) - Phis;
      dVbseffCV_dVb = -dPhis_dVb;
    }
    CoxWL = C.coxe * PP(weffCV) * PP(leffCV) * nf;
    if (C.cvchargeMod == 0) {
      noff = n * PP(noff);
      dnoff_dVd = PP(noff) * dn_dVd;
      dnoff_dVb = PP(noff) * dn_dVb;
      T0 = Vtm * noff;
      voffcv = PP(voffcv);
      VgstNVt = (Vgst - voffcv) / T0;
      if (VgstNVt > kExpThreshold) {
        Vgsteff = Vgst - voffcv;
        dVgsteff_dVg = dVgs_eff_dVg;
        dVgsteff_dVd = -dVth_dVd;
        dVgsteff_dVb = -dVth_dVb;
      } else if (VgstNVt < -kExpThreshold) {
        Vgsteff = T0 * log(1.0 + kMinExp);
        dVgsteff_dVg = 0.0;
        dVgsteff_dVd = Vgsteff / noff;
        dVgsteff_dVb = dVgsteff_dVd * dnoff_dVb;
        dVgsteff_dVd *= dnoff_dVd;
      } else {
        ExpVgst = exp(VgstNVt);
        Vgsteff = T0 * log(1.0 + ExpVgst);
        dVgsteff_dVg = ExpVgst / (1.0 + ExpVgst);
        dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd)
                       + Vgsteff / noff * dnoff_dVd;
        dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb)
                       + Vgsteff / noff * dnoff_dVb;
        dVgsteff_dVg *= dVgs_eff_dVg;
      }
    } else {
      T0 = n * Vtm;
      T1 = PP(mstarcv) * Vgst;
      T2 = T1 / T0;
      if (T2 > kExpThreshold) {
        T10 = T1;
        dT10_dVg = PP(mstarcv) * dVgs_eff_dVg;
        dT10_dVd = -dVth_dVd * PP(mstarcv);
        dT10_dVb = -dVth_dVb * PP(mstarcv);
      } else if (T2 < -kExpThreshold) {
        T10 = Vtm * log(1.0 + kMinExp);
        dT10_dVg = 0.0;
        dT10_dVd = T10 * dn_dVd;
        dT10_dVb = T10 * dn_dVb;
        T10 *= n;
      } else {
        ExpVgst = exp(T2);
        T3 = Vtm * log(1.0 + ExpVgst);
        T10 = n * T3;
        dT10_dVg = PP(mstarcv) * ExpVgst / (1.0 + ExpVgst);
        dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
        dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
        dT10_dVg *= dVgs_eff_dVg;
      }
      T1 = PP(voffcbncv) - (1.0 - PP(mstarcv)) * Vgst;
      T2 = T1 / T0;
      if (T2 < -kExpThreshold) {
        T3 = C.coxe * kMinExp / PP(cdep0);
        T9 = PP(mstarcv) + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
      } else if (T2 > kExpThreshold) {
        T3 = C.coxe * kMaxExp / PP(cdep0);
        T9 = PP(mstarcv) + T3 * n;
        dT9_dVg = 0.0;
        dT9_dVd = dn_dVd * T3;
        dT9_dVb = dn_dVb * T3;
      } else {
        ExpVgst = exp(T2);
        T3 = C.coxe / PP(cdep0);
        T4 = T3 * ExpVgst;
        T5 = T1 * T4 / T0;
        T9 = PP(mstarcv) + n * T4;
        dT9_dVg = T3 * (PP(mstarcv) - 1.0) * ExpVgst / Vtm;
        dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
        dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
        dT9_dVg *= dVgs_eff_dVg;
      }
      Vgsteff = T10 / T9;
      T11 = T9 * T9;
      dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
      dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
      dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;
    }
    if (X && capMod == 1) {
    double Vfb;
    /* capMod = 1 (:11139-11319) */
    Vfb = IV(vfbzb);
    V3 = Vfb - Vgs_eff + VbseffCV - kDelta3;
    if (Vfb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * kDelta3 * Vfb);
    else T0 = sqrt(V3 * V3 + 4.0 * kDelta3 * Vfb);
    T1 = 0.5 * (1.0 + V3 / T0);
    Vfbeff = Vfb - 0.5 * (V3 + T0);
    dVfbeff_dVg = T1 * dVgs_eff_dVg;
    dVfbeff_dVb = -T1 * dVbseffCV_dVb;
    Qac0 = CoxWL * (Vfbeff - Vfb);
    dQac0_dVg = CoxWL * dVfbeff_dVg;
    dQac0_dVb = CoxWL * dVfbeff_dVb;
    T0 = 0.5 * PP(k1ox);
    T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteff;
    if (PP(k1ox) == 0.0) {
      T1 = 0.0;
      T2 = 0.0;
    } else if (T3 < 0.0) {
      T1 = T0 + T3 / PP(k1ox);
      T2 = CoxWL;
    } else {
      T1 = sqrt(T0 * T0 + T3);
      T2 = CoxWL * T0 / T1;
    }
    Qsub0 = CoxWL * PP(k1ox) * (T1 - T0);
    dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg);
    dQsub0_dVd = -T2 * dVgsteff_dVd;
    dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb);
    AbulkCV = Abulk0 * PP(abulkCVfactor);
    dAbulkCV_dVb = PP(abulkCVfactor) * dAbulk0_dVb;
    VdsatCV = Vgsteff / AbulkCV;
    T0 = VdsatCV - Vds - kDelta4;
    dT0_dVg = 1.0 / AbulkCV;
    dT0_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
    T1 = sqrt(T0 * T0 + 4.0 * kDelta4 * VdsatCV);
    dT1_dVg = (T0 + kDelta4 + kDelta4) / T1;
    dT1_dVd = -T0 / T1;
    dT1_dVb = dT1_dVg * dT0_dVb;
    dT1_dVg *= dT0_dVg;
    if (T0 >= 0.0) {
      VdseffCV = VdsatCV - 0.5 * (T0 + T1);
      dVdseffCV_dVg = 0.5 * (dT0_dVg - dT1_dVg);
      dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
      dVdseffCV_dVb = 0.5 * (dT0_dVb - dT1_dVb);
    } else {
      T3 = (kDelta4 + kDelta4) / (T1 - T0);
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
    T1 = 12.0 * (Vgsteff - 0.5 * T0 + 1.0e-20);
    T2 = VdseffCV / T1;
    T3 = T0 * T2;
    T4 = (1.0 - 12.0 * T2 * T2 * AbulkCV);
    T5 = (6.0 * T0 * (4.0 * Vgsteff - T0) / (T1 * T1) - 0.5);
    T6 = 12.0 * T2 * T2 * Vgsteff;
    qgate = CoxWL * (Vgsteff - 0.5 * VdseffCV + T3);
    Cgg1 = CoxWL * (T4 + T5 * dVdseffCV_dVg);
    Cgd1 = CoxWL * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd;
    Cgb1 = CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteff_dVb;
    Cgg1 *= dVgsteff_dVg;
    T7 = 1.0 - AbulkCV;
    qbulk = CoxWL * T7 * (0.5 * VdseffCV - T3);
    T4 = -T7 * (T4 - 1.0);
    T5 = -T7 * T5;
    T6 = -(T7 * T6 + (0.5 * VdseffCV - T3));
    Cbg1 = CoxWL * (T4 + T5 * dVdseffCV_dVg);
    Cbd1 = CoxWL * T5 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd;
    Cbb1 = CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb;
    Cbg1 *= dVgsteff_dVg;
    if (C.xpart > 0.5) {
      T1 = T1 + T1;
      qsrc = -CoxWL * (0.5 * Vgsteff + 0.25 * T0 - T0 * T0 / T1);
      T7 = (4.0 * Vgsteff - T0) / (T1 * T1);
      T4 = -(0.5 + 24.0 * T0 * T0 / (T1 * T1));
      T5 = -(0.25 * AbulkCV - 12.0 * AbulkCV * T0 * T7);
      T6 = -(0.25 * VdseffCV - 12.0 * T0 * VdseffCV * T7);
      Csg = CoxWL * (T4 + T5 * dVdseffCV_dVg);
      Csd = CoxWL * T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd;
      Csb = CoxWL * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb;
      Csg *= dVgsteff_dVg;
    } else if (C.xpart < 0.5) {
      T1 = T1 / 12.0;
      T2 = 0.5 * CoxWL / (T1 * T1);
      T3 = Vgsteff * (2.0 * T0 * T0 / 3.0 + Vgsteff * (Vgsteff - 4.0 * T0 / 3.0))
           - 2.0 * T0 * T0 * T0 / 15.0;
      qsrc = -T2 * T3;
      T7 = 4.0 / 3.0 * Vgsteff * (Vgsteff - T0) + 0.4 * T0 * T0;
      T4 = -2.0 * qsrc / T1 - T2 * (Vgsteff * (3.0 * Vgsteff - 8.0 * T0 / 3.0)
                                       + 2.0 * T0 * T0 / 3.0);
      T5 = (qsrc / T1 + T2 * T7) * AbulkCV;
      T6 = (qsrc / T1 * VdseffCV + T2 * T7 * VdseffCV);
      Csg = (T4 + T5 * dVdseffCV_dVg);
      Csd = T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd;
      Csb = (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb;
      Csg *= dVgsteff_dVg;
    } else {
      qsrc = -0.5 * (qgate + qbulk);
      Csg = -0.5 * (Cgg1 + Cbg1);
      Csb = -0.5 * (Cgb1 + Cbb1);
      Csd = -0.5 * (Cgd1 + Cbd1);
    }
    qgate += Qac0 + Qsub0;
    qbulk -= (Qac0 + Qsub0);
    qdrn = -(qgate + qbulk + qsrc);
    Cgg = dQac0_dVg + dQsub0_dVg + Cgg1;
    Cgd = dQsub0_dVd + Cgd1;
    Cgb = dQac0_dVb + dQsub0_dVb + Cgb1;
    Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
    Cbd = Cbd1 - dQsub0_dVd;
    Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
    Cgb *= dVbseff_dVb;
    Cbb *= dVbseff_dVb;
    Csb *= dVbseff_dVb;
    cggb = Cgg;
    cgsb = -(Cgg + Cgd + Cgb);
    cgdb = Cgd;
    cdgb = -(Cgg + Cbg + Csg);
    cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
    cddb = -(Cgd + Cbd + Csd);
    cbgb = Cbg;
    cbsb = -(Cbg + Cbd + Cbb);
    cbdb = Cbd;
    } else {
    /* capMod = 2 (:11321-11602) */
    V3 = IV(vfbzb) - Vgs_eff + VbseffCV - kDelta3;
    if (IV(vfbzb) <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * kDelta3 * IV(vfbzb));
    else T0 = sqrt(V3 * V3 + 4.0 * kDelta3 * IV(vfbzb));
    T1 = 0.5 * (1.0 + V3 / T0);
    Vfbeff = IV(vfbzb) - 0.5 * (V3 + T0);
    dVfbeff_dVg = T1 * dVgs_eff_dVg;
    dVfbeff_dVb = -T1 * dVbseffCV_dVb;
    Cox = C.coxp;
    Tox = 1.0e8 * C.toxp;
    T0 = (Vgs_eff - VbseffCV - IV(vfbzb)) / Tox;
    dT0_dVg = dVgs_eff_dVg / Tox;
    dT0_dVb = -dVbseffCV_dVb / Tox;
    tmp = T0 * PP(acde);
    if ((-kExpThreshold < tmp) && (tmp < kExpThreshold)) {
      Tcen = PP(ldeb) * exp(tmp);
      dTcen_dVg = PP(acde) * Tcen;
      dTcen_dVb = dTcen_dVg * dT0_dVb;
      dTcen_dVg *= dT0_dVg;
    } else if (tmp <= -kExpThreshold) {
      Tcen = PP(ldeb) * kMinExp;
      dTcen_dVg = dTcen_dVb = 0.0;
    } else {
      Tcen = PP(ldeb) * kMaxExp;
      dTcen_dVg = dTcen_dVb = 0.0;
    }
    LINK = 1.0e-3 * C.toxp;
    V3 = PP(ldeb) - Tcen - LINK;
    V4 = sqrt(V3 * V3 + 4.0 * LINK * PP(ldeb));
    Tcen = PP(ldeb) - 0.5 * (V3 + V4);
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
    CoxWLcen = CoxWL * Coxeff / C.coxe;
    Qac0 = CoxWLcen * (Vfbeff - IV(vfbzb));
    QovCox = Qac0 / Coxeff;
    dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
    dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
    T0 = 0.5 * PP(k1ox);
    T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteff;
    if (PP(k1ox) == 0.0) {
      T1 = 0.0;
      T2 = 0.0;
    } else if (T3 < 0.0) {
      T1 = T0 + T3 / PP(k1ox);
      T2 = CoxWLcen;
    } else {
      T1 = sqrt(T0 * T0 + T3);
      T2 = CoxWLcen * T0 / T1;
    }
    Qsub0 = CoxWLcen * PP(k1ox) * (T1 - T0);
    QovCox = Qsub0 / Coxeff;
    dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg) + QovCox * dCoxeff_dVg;
    dQsub0_dVd = -T2 * dVgsteff_dVd;
    dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb) + QovCox * dCoxeff_dVb;
    if (PP(k1ox) <= 0.0) {
      Denomi = 0.25 * PP(moin) * Vtm;
      T0 = 0.5 * PP(sqrtPhi);
    } else {
      Denomi = PP(moin) * Vtm * PP(k1ox) * PP(k1ox);
      T0 = PP(k1ox) * PP(sqrtPhi);
    }
    T1 = 2.0 * T0 + Vgsteff;
    DeltaPhi = Vtm * log(1.0 + T1 * Vgsteff / Denomi);
    dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (Denomi + T1 * Vgsteff);
    T0 = Vgsteff - DeltaPhi - 0.001;
    dT0_dVg = 1.0 - dDeltaPhi_dVg;
    T1 = sqrt(T0 * T0 + Vgsteff * 0.004);
    VgDP = 0.5 * (T0 + T1);
    dVgDP_dVg = 0.5 * (dT0_dVg + (T0 * dT0_dVg + 0.002) / T1);
    Tox += Tox;
    T0 = (Vgsteff + IV(vtfbphi2)) / Tox;
    tmp = exp(C.bdos * 0.7 * log(T0));
    T1 = 1.0 + tmp;
    T2 = C.bdos * 0.7 * tmp / (T0 * Tox);
    Tcen = C.ados * 1.9e-9 / T1;
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
    CoxWLcen = CoxWL * Coxeff / C.coxe;
    AbulkCV = Abulk0 * PP(abulkCVfactor);
    dAbulkCV_dVb = PP(abulkCVfactor) * dAbulk0_dVb;
    VdsatCV = VgDP / AbulkCV;
    T0 = VdsatCV - Vds - kDelta4;
    dT0_dVg = dVgDP_dVg / AbulkCV;
    dT0_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
    T1 = sqrt(T0 * T0 + 4.0 * kDelta4 * VdsatCV);
    dT1_dVg = (T0 + kDelta4 + kDelta4) / T1;
    dT1_dVd = -T0 / T1;
    dT1_dVb = dT1_dVg * dT0_dVb;
    dT1_dVg *= dT0_dVg;
    if (T0 >= 0.0) {
      VdseffCV = VdsatCV - 0.5 * (T0 + T1);
      dVdseffCV_dVg = 0.5 * (dT0_dVg - dT1_dVg);
      dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
      dVdseffCV_dVb = 0.5 * (dT0_dVb - dT1_dVb);
    } else {
      T3 = (kDelta4 + kDelta4) / (T1 - T0);
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
    qgate = CoxWLcen * (T1 - T0 * (0.5 - T3));
    QovCox = qgate / Coxeff;
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
    qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
    QovCox = qbulk / Coxeff;
    Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
    Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
    Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
    Cbg1 = Cbg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
    if (C.xpart > 0.5) {
      qsrc = -CoxWLcen * (T1 / 2.0 + T0 / 4.0 - 0.5 * T0 * T0 / T2);
      QovCox = qsrc / Coxeff;
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
    } else if (C.xpart < 0.5) {
      T2 = T2 / 12.0;
      T3 = 0.5 * CoxWLcen / (T2 * T2);
      T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
      qsrc = -T3 * T4;
      QovCox = qsrc / Coxeff;
      T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
      T5 = -2.0 * qsrc / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
      T6 = AbulkCV * (qsrc / T2 + T3 * T8);
      T7 = T6 * VdseffCV / AbulkCV;
      Csg = T5 * dVgDP_dVg + T6 * dVdseffCV_dVg;
      Csd = Csg * dVgsteff_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
      Csb = Csg * dVgsteff_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
      Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
    } else {
      qsrc = -0.5 * qgate;
      Csg = -0.5 * Cgg1;
      Csd = -0.5 * Cgd1;
      Csb = -0.5 * Cgb1;
    }
    qgate += Qac0 + Qsub0 - qbulk;
    qbulk -= (Qac0 + Qsub0);
    qdrn = -(qgate + qbulk + qsrc);
    Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
    Cbd = Cbd1 - dQsub0_dVd;
    Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
    Cgg = Cgg1 - Cbg;
    Cgd = Cgd1 - Cbd;
    Cgb = Cgb1 - Cbb;
    Cgb *= dVbseff_dVb;
    Cbb *= dVbseff_dVb;
    Csb *= dVbseff_dVb;
    cggb = Cgg;
    cgsb = -(Cgg + Cgd + Cgb);
    cgdb = Cgd;
    cdgb = -(Cgg + Cbg + Csg);
    cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
    cddb = -(Cgd + Cbd + Csd);
    cbgb = Cbg;
    cbsb = -(Cbg + Cbd + Cbb);
    cbdb = Cbd;
    }
  }
  csgb = -cggb - cdgb - cbgb;
  csdb = -cgdb - cddb - cbdb;
  cssb = -cgsb - cdsb - cbsb;
  cgbb = -cgdb - cggb - cgsb;
  cdbb = -cddb - cdgb - cdsb;
  cbbb = -cbgb - cbdb - cbsb;
  csbb = -cgbb - cdbb - cbbb;

  if (RELOAD) reload_params(dfb);
  /* NQS (:11626-11660) */
  if (X && ((itrnqsMod) || (iacnqsMod))) {
    qchqs = qcheq = -(qbulk + qgate);
    cqgb = -(cggb + cbgb);
    cqdb = -(cgdb + cbdb);
    cqsb = -(cgsb + cbsb);
    cqbb = -(cqgb + cqdb + cqsb);
    CoxWL = C.coxe * PP(weffCV) * nf * PP(leffCV);
    T1 = gcrg / CoxWL;
    gtau = T1 * kScalingFactor;
    if (iacnqsMod) taunet = 1.0 / T1;
  }

  /* junction charges (:11662-11810) */
  if (ChargeComputationNeeded) {
    czbd = C.DunitAreaTempJctCap * IV(Adeff);
    czbs = C.SunitAreaTempJctCap * IV(Aseff);
    czbdsw = C.DunitLengthSidewallTempJctCap * IV(Pdeff);
    czbdswg = C.DunitLengthGateSidewallTempJctCap * PP(weffCJ) * nf;
    czbssw = C.SunitLengthSidewallTempJctCap * IV(Pseff);
    czbsswg = C.SunitLengthGateSidewallTempJctCap * PP(weffCJ) * nf;
    MJS = C.SbulkJctBotGradingCoeff;
    MJSWS = C.SbulkJctSideGradingCoeff;
    MJSWGS = C.SbulkJctGateSideGradingCoeff;
    MJD = C.DbulkJctBotGradingCoeff;
    MJSWD = C.DbulkJctSideGradingCoeff;
    MJSWGD = C.DbulkJctGateSideGradingCoeff;
    if (vbs_jct == 0.0) {
      qbs = 0.0;
      capbs = czbs + czbssw + czbsswg;
    } else if (vbs_jct < 0.0) {
      if (czbs > 0.0) {
        arg = 1.0 - vbs_jct / C.PhiBS;
        if (MJS == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJS * log(arg));
        qbs = C.PhiBS * czbs * (1.0 - arg * sarg) / (1.0 - MJS);
        capbs = czbs * sarg;
      } else {
        qbs = 0.0;
        capbs = 0.0;
      }
      if (czbssw > 0.0) {
        arg = 1.0 - vbs_jct / C.PhiBSWS;
        if (MJSWS == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWS * log(arg));
        qbs += C.PhiBSWS * czbssw * (1.0 - arg * sarg) / (1.0 - MJSWS);
        capbs += czbssw * sarg;
      }
      if (czbsswg > 0.0) {
        arg = 1.0 - vbs_jct / C.PhiBSWGS;
        if (MJSWGS == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWGS * log(arg));
        qbs += C.PhiBSWGS * czbsswg * (1.0 - arg * sarg) / (1.0 - MJSWGS);
        capbs += czbsswg * sarg;
      }
    } else {
      T0 = czbs + czbssw + czbsswg;
      T1 = vbs_jct * (czbs * MJS / C.PhiBS + czbssw * MJSWS / C.PhiBSWS + czbsswg * MJSWGS / C.PhiBSWGS);
      qbs = vbs_jct * (T0 + 0.5 * T1);
      capbs = T0 + T1;
    }
    if (vbd_jct == 0.0) {
      qbd = 0.0;
      capbd = czbd + czbdsw + czbdswg;
    } else if (vbd_jct < 0.0) {
      if (czbd > 0.0) {
        arg = 1.0 - vbd_jct / C.PhiBD;
        if (MJD == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJD * log(arg));
        qbd = C.PhiBD * czbd * (1.0 - arg * sarg) / (1.0 - MJD);
        capbd = czbd * sarg;
      } else {
        qbd = 0.0;
        capbd = 0.0;
      }
      if (czbdsw > 0.0) {
        arg = 1.0 - vbd_jct / C.PhiBSWD;
        if (MJSWD == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWD * log(arg));
        qbd += C.PhiBSWD * czbdsw * (1.0 - arg * sarg) / (1.0 - MJSWD);
        capbd += czbdsw * sarg;
      }
      if (czbdswg > 0.0) {
        arg = 1.0 - vbd_jct / C.PhiBSWGD;
        if (MJSWGD == 0.5) sarg = 1.0 / sqrt(arg);
        else sarg = exp(-MJSWGD * log(arg));
        qbd += C.PhiBSWGD * czbdswg * (1.0 - arg * sarg) / (1.0 - MJSWGD);
        capbd += czbdswg * sarg;
      }
    } else {
      T0 = czbd + czbdsw + czbdswg;
      T1 = vbd_jct * (czbd * MJD / C.PhiBD + czbdsw * MJSWD / C.PhiBSWD + czbdswg * MJSWGD / C.PhiBSWGD);
      qbd = vbd_jct * (T0 + 0.5 * T1);
      capbd = T0 + T1;
    }
  }

  if (RELOAD) reload_params(dfb);
  if constexpr (X) {
    /* gate resistor currents and overlap capacitances (:11812-11883) */
    if (irgateMod == 3) {
      vgdx = vgmd;
      vgsx = vgms;
    } else {
      vgdx = vgd;
      vgsx = vgs;
    }
    Igate = IgateMid = 0.0;
    Igate_Jdxp = IgateMid_Jdxp = 0.0;
    if (irgateMod == 1) {
      Igate = XP(grgeltd) * (Vgegp);
    } else if (irgateMod == 2) {
      Igate = (gcrg) * (Vgegp);
    } else if (irgateMod == 3) {
      Igate = XP(grgeltd) * (Vgegm);
      IgateMid = gcrg * (Vgmgp);
    }
    if (capMod == 0) {
      cgdo = PP(cgdo);
      qgdo = PP(cgdo) * vgdx;
      cgso = PP(cgso);
      qgso = PP(cgso) * vgsx;
    } else {
    T0 = vgdx + kDelta1;
    T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
    T2 = 0.5 * (T0 - T1);
    T3 = PP(weffCV) * PP(cgdl);
    T4 = sqrt(1.0 - 4.0 * T2 / PP(ckappad));
    cgdo = PP(cgdo) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
    qgdo = (PP(cgdo) + T3) * vgdx - T3 * (T2 + 0.5 * PP(ckappad) * (T4 - 1.0));
    T0 = vgsx + kDelta1;
    T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
    T2 = 0.5 * (T0 - T1);
    T3 = PP(weffCV) * PP(cgsl);
    T4 = sqrt(1.0 - 4.0 * T2 / PP(ckappas));
    cgso = PP(cgso) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
    qgso = (PP(cgso) + T3) * vgsx - T3 * (T2 + 0.5 * PP(ckappas) * (T4 - 1.0));
    }
    if (nf != 1.0) {
      cgdo *= nf;
      cgso *= nf;
      qgdo *= nf;
      qgso *= nf;
    }
  } else {
    /* overlap capacitances, capMod != 0 (:11812-11883) */
    vgdx = vgd;
    vgsx = vgs;
    Igate = IgateMid = 0.0;
    Igate_Jdxp = IgateMid_Jdxp = 0.0;
    T0 = vgdx + kDelta1;
    T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
    T2 = 0.5 * (T0 - T1);
    T3 = PP(weffCV) * PP(cgdl);
    T4 = sqrt(1.0 - 4.0 * T2 / PP(ckappad));
    cgdo = PP(cgdo) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
    qgdo = (PP(cgdo) + T3) * vgdx - T3 * (T2 + 0.5 * PP(ckappad) * (T4 - 1.0));
    T0 = vgsx + kDelta1;
    T1 = sqrt(T0 * T0 + 4.0 * kDelta1);
    T2 = 0.5 * (T0 - T1);
    T3 = PP(weffCV) * PP(cgsl);
    T4 = sqrt(1.0 - 4.0 * T2 / PP(ckappas));
    cgso = PP(cgso) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
    qgso = (PP(cgso) + T3) * vgsx - T3 * (T2 + 0.5 * PP(ckappas) * (T4 - 1.0));
    if (nf != 1.0) {
      cgdo *= nf;
      cgso *= nf;
      qgdo *= nf;
      qgso *= nf;
    }
  }

  // ---- setupCapacitors_oldDAE / setupCapacitors_newDAE ----
  if constexpr (X) {
    {

    double ag0 = ss.pdt, Cdd, Csd, Cdg, Csg, Cds, Css;
    if (!(ss.dcopFlag) && ss.initTranFlag && ss.newtonIter == 0) ag0 = 0.0;
    if (mode > 0) {
      if (itrnqsMod == 0) {
        qdrn -= qgdo;
        if (irgateMod == 3) {
          gcgmgmb = (cgdo + cgso + PP(cgbo)) * ag0;
          gcgmdb = -cgdo * ag0;
          gcgmsb = -cgso * ag0;
          gcgmbb = -PP(cgbo) * ag0;
          gcdgmb = gcgmdb;
          gcsgmb = gcgmsb;
          gcbgmb = gcgmbb;
          gcggb = cggb * ag0;
          gcgdb = cgdb * ag0;
          gcgsb = cgsb * ag0;
          gcgbb = -(gcggb + gcgdb + gcgsb);
          gcdgb = cdgb * ag0;
          gcsgb = -(cggb + cbgb + cdgb) * ag0;
          gcbgb = cbgb * ag0;
          qgmb = PP(cgbo) * vgmb;
          qgmid = qgdo + qgso + qgmb;
          qbulk -= qgmb;
          qsrc = -(qgate + qgmid + qbulk + qdrn);
        } else {
          gcggb = (cggb + cgdo + cgso + PP(cgbo)) * ag0;
          gcgdb = (cgdb - cgdo) * ag0;
          gcgsb = (cgsb - cgso) * ag0;
          gcgbb = -(gcggb + gcgdb + gcgsb);
          gcdgb = (cdgb - cgdo) * ag0;
          gcsgb = -(cggb + cbgb + cdgb + cgso) * ag0;
          gcbgb = (cbgb - PP(cgbo)) * ag0;
          gcdgmb = gcsgmb = gcbgmb = 0.0;
          qgb = PP(cgbo) * vgb;
          qgate += qgdo + qgso + qgb;
          qbulk -= qgb;
          qsrc = -(qgate + qbulk + qdrn);
        }
        gcddb = (cddb + capbd + cgdo) * ag0;
        gcdsb = cdsb * ag0;
        gcsdb = -(cgdb + cbdb + cddb) * ag0;
        gcssb = (capbs + cgso - (cgsb + cbsb + cdsb)) * ag0;
        if (!irbodyMod) {
          gcdbb = -(gcdgb + gcddb + gcdsb + gcdgmb);
          gcsbb = -(gcsgb + gcsdb + gcssb + gcsgmb);
          gcbdb = (cbdb - capbd) * ag0;
          gcbsb = (cbsb - capbs) * ag0;
          gcdbdb = 0.0;
          gcsbsb = 0.0;
        } else {
          gcdbb = -(cddb + cdgb + cdsb) * ag0;
          gcsbb = -(gcsgb + gcsdb + gcssb + gcsgmb) + capbs * ag0;
          gcbdb = cbdb * ag0;
          gcbsb = cbsb * ag0;
          gcdbdb = -capbd * ag0;
          gcsbsb = -capbs * ag0;
        }
        gcbbb = -(gcbdb + gcbgb + gcbsb + gcbgmb);
        ggtg = ggtd = ggtb = ggts = 0.0;
        sxpart = 0.6;
        dxpart = 0.4;
        ddxpart_dVd = ddxpart_dVg = ddxpart_dVb = ddxpart_dVs = 0.0;
        dsxpart_dVd = dsxpart_dVg = dsxpart_dVb = dsxpart_dVs = 0.0;
      } else {
        qcheq = qchqs;
        CoxWL = C.coxe * PP(weffCV) * IV(nf) * PP(leffCV);
        T0 = qdef * kScalingFactor / CoxWL;
        ggtg = gtg = T0 * gcrgg;
        ggtd = gtd = T0 * gcrgd;
        ggts = gts = T0 * gcrgs;
        ggtb = gtb = T0 * gcrgb;
        gqdef = kScalingFactor * ag0;
        gcqgb = cqgb * ag0;
        gcqdb = cqdb * ag0;
        gcqsb = cqsb * ag0;
        gcqbb = cqbb * ag0;
        if (fabs(qcheq) <= 1.0e-5 * CoxWL) {
          if (C.xpart < 0.5) dxpart = 0.4;
          else if (C.xpart > 0.5) dxpart = 0.0;
          else dxpart = 0.5;
          ddxpart_dVd = ddxpart_dVg = ddxpart_dVb = ddxpart_dVs = 0.0;
        } else {
          dxpart = qdrn / qcheq;
          Cdd = cddb;
          Csd = -(cgdb + cddb + cbdb);
          ddxpart_dVd = (Cdd - dxpart * (Cdd + Csd)) / qcheq;
          Cdg = cdgb;
          Csg = -(cggb + cdgb + cbgb);
          ddxpart_dVg = (Cdg - dxpart * (Cdg + Csg)) / qcheq;
          Cds = cdsb;
          Css = -(cgsb + cdsb + cbsb);
          ddxpart_dVs = (Cds - dxpart * (Cds + Css)) / qcheq;
          ddxpart_dVb = -(ddxpart_dVd + ddxpart_dVg + ddxpart_dVs);
        }
        sxpart = 1.0 - dxpart;
        dsxpart_dVd = -ddxpart_dVd;
        dsxpart_dVg = -ddxpart_dVg;
        dsxpart_dVs = -ddxpart_dVs;
        dsxpart_dVb = -(dsxpart_dVd + dsxpart_dVg + dsxpart_dVs);
        if (irgateMod == 3) {
          gcgmgmb = (cgdo + cgso + PP(cgbo)) * ag0;
          gcgmdb = -cgdo * ag0;
          gcgmsb = -cgso * ag0;
          gcgmbb = -PP(cgbo) * ag0;
          gcdgmb = gcgmdb;
          gcsgmb = gcgmsb;
          gcbgmb = gcgmbb;
          gcdgb = gcsgb = gcbgb = 0.0;
          gcggb = gcgdb = gcgsb = gcgbb = 0.0;
          qgmb = PP(cgbo) * vgmb;
          qgmid = qgdo + qgso + qgmb;
          qgate = 0.0;
          qbulk = -qgmb;
          qdrn = -qgdo;
          qsrc = -(qgmid + qbulk + qdrn);
        } else {
          gcggb = (cgdo + cgso + PP(cgbo)) * ag0;
          gcgdb = -cgdo * ag0;
          gcgsb = -cgso * ag0;
          gcgbb = -PP(cgbo) * ag0;
          gcdgb = gcgdb;
          gcsgb = gcgsb;
          gcbgb = gcgbb;
          gcdgmb = gcsgmb = gcbgmb = 0.0;
          qgb = PP(cgbo) * vgb;
          qgate = qgdo + qgso + qgb;
          qbulk = -qgb;
          qdrn = -qgdo;
          qsrc = -(qgate + qbulk + qdrn);
        }
        gcddb = (capbd + cgdo) * ag0;
        gcdsb = gcsdb = 0.0;
        gcssb = (capbs + cgso) * ag0;
        if (!irbodyMod) {
          gcdbb = -(gcdgb + gcddb + gcdgmb);
          gcsbb = -(gcsgb + gcssb + gcsgmb);
          gcbdb = -capbd * ag0;
          gcbsb = -capbs * ag0;
          gcdbdb = 0.0;
          gcsbsb = 0.0;
        } else {
          gcdbb = gcsbb = gcbdb = gcbsb = 0.0;
          gcdbdb = -capbd * ag0;
          gcsbsb = -capbs * ag0;
        }
        gcbbb = -(gcbdb + gcbgb + gcbsb + gcbgmb);
      }
    } else {
      if (itrnqsMod == 0) {
        qsrc = qdrn - qgso;
        if (irgateMod == 3) {
          gcgmgmb = (cgdo + cgso + PP(cgbo)) * ag0;
          gcgmdb = -cgdo * ag0;
          gcgmsb = -cgso * ag0;
          gcgmbb = -PP(cgbo) * ag0;
          gcdgmb = gcgmdb;
          gcsgmb = gcgmsb;
          gcbgmb = gcgmbb;
          gcggb = cggb * ag0;
          gcgdb = cgsb * ag0;
          gcgsb = cgdb * ag0;
          gcgbb = -(gcggb + gcgdb + gcgsb);
          gcdgb = -(cggb + cbgb + cdgb) * ag0;
          gcsgb = cdgb * ag0;
          gcbgb = cbgb * ag0;
          qgmb = PP(cgbo) * vgmb;
          qgmid = qgdo + qgso + qgmb;
          qbulk -= qgmb;
          qdrn = -(qgate + qgmid + qbulk + qsrc);
        } else {
          gcggb = (cggb + cgdo + cgso + PP(cgbo)) * ag0;
          gcgdb = (cgsb - cgdo) * ag0;
          gcgsb = (cgdb - cgso) * ag0;
          gcgbb = -(gcggb + gcgdb + gcgsb);
          gcdgb = -(cggb + cbgb + cdgb + cgdo) * ag0;
          gcsgb = (cdgb - cgso) * ag0;
          gcbgb = (cbgb - PP(cgbo)) * ag0;
          gcdgmb = gcsgmb = gcbgmb = 0.0;
          qgb = PP(cgbo) * vgb;
          qgate += qgdo + qgso + qgb;
          qbulk -= qgb;
          qdrn = -(qgate + qbulk + qsrc);
        }
        gcddb = (capbd + cgdo - (cgsb + cbsb + cdsb)) * ag0;
        gcdsb = -(cgdb + cbdb + cddb) * ag0;
        gcsdb = cdsb * ag0;
        gcssb = (cddb + capbs + cgso) * ag0;
        if (!irbodyMod) {
          gcdbb = -(gcdgb + gcddb + gcdsb + gcdgmb);
          gcsbb = -(gcsgb + gcsdb + gcssb + gcsgmb);
          gcbdb = (cbsb - capbd) * ag0;
          gcbsb = (cbdb - capbs) * ag0;
          gcdbdb = 0.0;
          gcsbsb = 0.0;
        } else {
          gcdbb = -(gcdgb + gcddb + gcdsb + gcdgmb) + capbd * ag0;
          gcsbb = -(cddb + cdgb + cdsb) * ag0;
          gcbdb = cbsb * ag0;
          gcbsb = cbdb * ag0;
          gcdbdb = -capbd * ag0;
          gcsbsb = -capbs * ag0;
        }
        gcbbb = -(gcbgb + gcbdb + gcbsb + gcbgmb);
        ggtg = ggtd = ggtb = ggts = 0.0;
        sxpart = 0.4;
        dxpart = 0.6;
        ddxpart_dVd = ddxpart_dVg = ddxpart_dVb = ddxpart_dVs = 0.0;
        dsxpart_dVd = dsxpart_dVg = dsxpart_dVb = dsxpart_dVs = 0.0;
      } else {
        qcheq = qchqs;
        CoxWL = C.coxe * PP(weffCV) * IV(nf) * PP(leffCV);
        T0 = qdef * kScalingFactor / CoxWL;
        ggtg = gtg = T0 * gcrgg;
        ggts = gts = T0 * gcrgd;
        ggtd = gtd = T0 * gcrgs;
        ggtb = gtb = T0 * gcrgb;
        gqdef = kScalingFactor * ag0;
        gcqgb = cqgb * ag0;
        gcqdb = cqsb * ag0;
        gcqsb = cqdb * ag0;
        gcqbb = cqbb * ag0;
        if (fabs(qcheq) <= 1.0e-5 * CoxWL) {
          if (C.xpart < 0.5) sxpart = 0.4;
          else if (C.xpart > 0.5) sxpart = 0.0;
          else sxpart = 0.5;
          dsxpart_dVd = dsxpart_dVg = dsxpart_dVb = dsxpart_dVs = 0.0;
        } else {
          sxpart = qdrn / qcheq;
          Css = cddb;
          Cds = -(cgdb + cddb + cbdb);
          dsxpart_dVs = (Css - sxpart * (Css + Cds)) / qcheq;
          Csg = cdgb;
          Cdg = -(cggb + cdgb + cbgb);
          dsxpart_dVg = (Csg - sxpart * (Csg + Cdg)) / qcheq;
          Csd = cdsb;
          Cdd = -(cgsb + cdsb + cbsb);
          dsxpart_dVd = (Csd - sxpart * (Csd + Cdd)) / qcheq;
          dsxpart_dVb = -(dsxpart_dVd + dsxpart_dVg + dsxpart_dVs);
        }
        dxpart = 1.0 - sxpart;
        ddxpart_dVd = -dsxpart_dVd;
        ddxpart_dVg = -dsxpart_dVg;
        ddxpart_dVs = -dsxpart_dVs;
        ddxpart_dVb = -(ddxpart_dVd + ddxpart_dVg + ddxpart_dVs);
        if (irgateMod == 3) {
          gcgmgmb = (cgdo + cgso + PP(cgbo)) * ag0;
          gcgmdb = -cgdo * ag0;
          gcgmsb = -cgso * ag0;
          gcgmbb = -PP(cgbo) * ag0;
          gcdgmb = gcgmdb;
          gcsgmb = gcgmsb;
          gcbgmb = gcgmbb;
          gcdgb = gcsgb = gcbgb = 0.0;
          gcggb = gcgdb = gcgsb = gcgbb = 0.0;
          qgmb = PP(cgbo) * vgmb;
          qgmid = qgdo + qgso + qgmb;
          qgate = 0.0;
          qbulk = -qgmb;
          qdrn = -qgdo;
          qsrc = -qgso;
        } else {
          gcggb = (cgdo + cgso + PP(cgbo)) * ag0;
          gcgdb = -cgdo * ag0;
          gcgsb = -cgso * ag0;
          gcgbb = -PP(cgbo) * ag0;
          gcdgb = gcgdb;
          gcsgb = gcgsb;
          gcbgb = gcgbb;
          gcdgmb = gcsgmb = gcbgmb = 0.0;
          qgb = PP(cgbo) * vgb;
          qgate = qgdo + qgso + qgb;
          qbulk = -qgb;
          qdrn = -qgdo;
          qsrc = -qgso;
        }
        gcddb = (capbd + cgdo) * ag0;
        gcdsb = gcsdb = 0.0;
        gcssb = (capbs + cgso) * ag0;
        if (!irbodyMod) {
          gcdbb = -(gcdgb + gcddb + gcdgmb);
          gcsbb = -(gcsgb + gcssb + gcsgmb);
          gcbdb = -capbd * ag0;
          gcbsb = -capbs * ag0;
          gcdbdb = 0.0;
          gcsbsb = 0.0;
        } else {
          gcdbb = gcsbb = gcbdb = gcbsb = 0.0;
          gcdbdb = -capbd * ag0;
          gcsbsb = -capbs * ag0;
        }
        gcbbb = -(gcbdb + gcbgb + gcbsb + gcbgmb);
      }
    }
    }
    {

    if (mode > 0) {
      if (itrnqsMod == 0) {
        if (irgateMod == 3) {
          CAPcgmgmb = (cgdo + cgso + PP(cgbo));
          CAPcgmdb = -cgdo;
          CAPcgmsb = -cgso;
          CAPcgmbb = -PP(cgbo);
          CAPcdgmb = CAPcgmdb;
          CAPcsgmb = CAPcgmsb;
          CAPcbgmb = CAPcgmbb;
          CAPcggb = cggb;
          CAPcgdb = cgdb;
          CAPcgsb = cgsb;
          CAPcgbb = -(CAPcggb + CAPcgdb + CAPcgsb);
          CAPcdgb = cdgb;
          CAPcsgb = -(cggb + cbgb + cdgb);
          CAPcbgb = cbgb;
        } else {
          CAPcggb = (cggb + cgdo + cgso + PP(cgbo));
          CAPcgdb = (cgdb - cgdo);
          CAPcgsb = (cgsb - cgso);
          CAPcgbb = -(CAPcggb + CAPcgdb + CAPcgsb);
          CAPcdgb = (cdgb - cgdo);
          CAPcsgb = -(cggb + cbgb + cdgb + cgso);
          CAPcbgb = (cbgb - PP(cgbo));
          CAPcdgmb = CAPcsgmb = CAPcbgmb = 0.0;
        }
        CAPcddb = (cddb + capbd + cgdo);
        CAPcdsb = cdsb;
        CAPcsdb = -(cgdb + cbdb + cddb);
        CAPcssb = (capbs + cgso - (cgsb + cbsb + cdsb));
        if (!irbodyMod) {
          CAPcdbb = -(CAPcdgb + CAPcddb + CAPcdsb + CAPcdgmb);
          CAPcsbb = -(CAPcsgb + CAPcsdb + CAPcssb + CAPcsgmb);
          CAPcbdb = (cbdb - capbd);
          CAPcbsb = (cbsb - capbs);
          CAPcdbdb = 0.0;
          CAPcsbsb = 0.0;
        } else {
          CAPcdbb = -(cddb + cdgb + cdsb);
          CAPcsbb = -(CAPcsgb + CAPcsdb + CAPcssb + CAPcsgmb) + capbs;
          CAPcbdb = cbdb;
          CAPcbsb = cbsb;
          CAPcdbdb = -capbd;
          CAPcsbsb = -capbs;
        }
        CAPcbbb = -(CAPcbdb + CAPcbgb + CAPcbsb + CAPcbgmb);
      } else {
        CAPcqgb = cqgb;
        CAPcqdb = cqdb;
        CAPcqsb = cqsb;
        CAPcqbb = cqbb;
        if (irgateMod == 3) {
          CAPcgmgmb = (cgdo + cgso + PP(cgbo));
          CAPcgmdb = -cgdo;
          CAPcgmsb = -cgso;
          CAPcgmbb = -PP(cgbo);
          CAPcdgmb = CAPcgmdb;
          CAPcsgmb = CAPcgmsb;
          CAPcbgmb = CAPcgmbb;
          CAPcdgb = CAPcsgb = CAPcbgb = 0.0;
          CAPcggb = CAPcgdb = CAPcgsb = CAPcgbb = 0.0;
        } else {
          CAPcggb = (cgdo + cgso + PP(cgbo));
          CAPcgdb = -cgdo;
          CAPcgsb = -cgso;
          CAPcgbb = -PP(cgbo);
          CAPcdgb = CAPcgdb;
          CAPcsgb = CAPcgsb;
          CAPcbgb = CAPcgbb;
          CAPcdgmb = CAPcsgmb = CAPcbgmb = 0.0;
        }
        CAPcddb = (capbd + cgdo);
        CAPcdsb = CAPcsdb = 0.0;
        CAPcssb = (capbs + cgso);
        if (!irbodyMod) {
          CAPcdbb = -(CAPcdgb + CAPcddb + CAPcdgmb);
          CAPcsbb = -(CAPcsgb + CAPcssb + CAPcsgmb);
          CAPcbdb = -capbd;
          CAPcbsb = -capbs;
          CAPcdbdb = 0.0;
          CAPcsbsb = 0.0;
        } else {
          CAPcdbb = CAPcsbb = CAPcbdb = CAPcbsb = 0.0;
          CAPcdbdb = -capbd;
          CAPcsbsb = -capbs;
        }
        CAPcbbb = -(CAPcbdb + CAPcbgb + CAPcbsb + CAPcbgmb);
      }
    } else {
      if (itrnqsMod == 0) {
        if (irgateMod == 3) {
          CAPcgmgmb = (cgdo + cgso + PP(cgbo));
          CAPcgmdb = -cgdo;
          CAPcgmsb = -cgso;
          CAPcgmbb = -PP(cgbo);
          CAPcdgmb = CAPcgmdb;
          CAPcsgmb = CAPcgmsb;
          CAPcbgmb = CAPcgmbb;
          CAPcggb = cggb;
          CAPcgdb = cgsb;
          CAPcgsb = cgdb;
          CAPcgbb = -(CAPcggb + CAPcgdb + CAPcgsb);
          CAPcdgb = -(cggb + cbgb + cdgb);
          CAPcsgb = cdgb;
          CAPcbgb = cbgb;
        } else {
          CAPcggb = (cggb + cgdo + cgso + PP(cgbo));
          CAPcgdb = (cgsb - cgdo);
          CAPcgsb = (cgdb - cgso);
          CAPcgbb = -(CAPcggb + CAPcgdb + CAPcgsb);
          CAPcdgb = -(cggb + cbgb + cdgb + cgdo);
          CAPcsgb = (cdgb - cgso);
          CAPcbgb = (cbgb - PP(cgbo));
          CAPcdgmb = CAPcsgmb = CAPcbgmb = 0.0;
        }
        CAPcddb = (capbd + cgdo - (cgsb + cbsb + cdsb));
        CAPcdsb = -(cgdb + cbdb + cddb);
        CAPcsdb = cdsb;
        CAPcssb = (cddb + capbs + cgso);
        if (!irbodyMod) {
          CAPcdbb = -(CAPcdgb + CAPcddb + CAPcdsb + CAPcdgmb);
          CAPcsbb = -(CAPcsgb + CAPcsdb + CAPcssb + CAPcsgmb);
          CAPcbdb = (cbsb - capbd);
          CAPcbsb = (cbdb - capbs);
          CAPcdbdb = 0.0;
          CAPcsbsb = 0.0;
        } else {
          CAPcdbb = -(CAPcdgb + CAPcddb + CAPcdsb + CAPcdgmb) + capbd;
          CAPcsbb = -(cddb + cdgb + cdsb);
          CAPcbdb = cbsb;
          CAPcbsb = cbdb;
          CAPcdbdb = -capbd;
          CAPcsbsb = -capbs;
        }
        CAPcbbb = -(CAPcbgb + CAPcbdb + CAPcbsb + CAPcbgmb);
      } else {
        CAPcqgb = cqgb;
        CAPcqdb = cqsb;
        CAPcqsb = cqdb;
        CAPcqbb = cqbb;
        if (irgateMod == 3) {
          CAPcgmgmb = (cgdo + cgso + PP(cgbo));
          CAPcgmdb = -cgdo;
          CAPcgmsb = -cgso;
          CAPcgmbb = -PP(cgbo);
          CAPcdgmb = CAPcgmdb;
          CAPcsgmb = CAPcgmsb;
          CAPcbgmb = CAPcgmbb;
          CAPcdgb = CAPcsgb = CAPcbgb = 0.0;
          CAPcggb = CAPcgdb = CAPcgsb = CAPcgbb = 0.0;
        } else {
          CAPcggb = (cgdo + cgso + PP(cgbo));
          CAPcgdb = -cgdo;
          CAPcgsb = -cgso;
          CAPcgbb = -PP(cgbo);
          CAPcdgb = CAPcgdb;
          CAPcsgb = CAPcgsb;
          CAPcbgb = CAPcgbb;
          CAPcdgmb = CAPcsgmb = CAPcbgmb = 0.0;
        }
        CAPcddb = (capbd + cgdo);
        CAPcdsb = CAPcsdb = 0.0;
        CAPcssb = (capbs + cgso);
        if (!irbodyMod) {
          CAPcdbb = -(CAPcdgb + CAPcddb + CAPcdgmb);
          CAPcsbb = -(CAPcsgb + CAPcssb + CAPcsgmb);
          CAPcbdb = -capbd;
          CAPcbsb = -capbs;
          CAPcdbdb = 0.0;
          CAPcsbsb = 0.0;
        } else {
          CAPcdbb = CAPcsbb = CAPcbdb = CAPcbsb = 0.0;
          CAPcdbdb = -capbd;
          CAPcsbsb = -capbs;
        }
        CAPcbbb = -(CAPcbdb + CAPcbgb + CAPcbsb + CAPcbgmb);
      }
    }
    }
  } else {
    {
      double ag0 = ss.pdt;
      if (!(ss.dcopFlag) && ss.initTranFlag && ss.newtonIter == 0) ag0 = 0.0;
      if (mode > 0) {
        qdrn -= qgdo;
        gcggb = (cggb + cgdo + cgso + PP(cgbo)) * ag0;
        gcgdb = (cgdb - cgdo) * ag0;
        gcgsb = (cgsb - cgso) * ag0;
        gcgbb = -(gcggb + gcgdb + gcgsb);
        gcdgb = (cdgb - cgdo) * ag0;
        gcsgb = -(cggb + cbgb + cdgb + cgso) * ag0;
        gcbgb = (cbgb - PP(cgbo)) * ag0;
        gcdgmb = gcsgmb = gcbgmb = 0.0;
        qgb = PP(cgbo) * vgb;
        qgate += qgdo + qgso + qgb;
        qbulk -= qgb;
        qsrc = -(qgate + qbulk + qdrn);
        gcddb = (cddb + capbd + cgdo) * ag0;
        gcdsb = cdsb * ag0;
        gcsdb = -(cgdb + cbdb + cddb) * ag0;
        gcssb = (capbs + cgso - (cgsb + cbsb + cdsb)) * ag0;
        gcdbb = -(gcdgb + gcddb + gcdsb + gcdgmb);
        gcsbb = -(gcsgb + gcsdb + gcssb + gcsgmb);
        gcbdb = (cbdb - capbd) * ag0;
        gcbsb = (cbsb - capbs) * ag0;
        gcdbdb = 0.0;
        gcsbsb = 0.0;
        gcbbb = -(gcbdb + gcbgb + gcbsb + gcbgmb);
        ggtg = ggtd = ggtb = ggts = 0.0;
        sxpart = 0.6;
        dxpart = 0.4;
        ddxpart_dVd = ddxpart_dVg = ddxpart_dVb = ddxpart_dVs = 0.0;
        dsxpart_dVd = dsxpart_dVg = dsxpart_dVb = dsxpart_dVs = 0.0;
      } else {
        qsrc = qdrn - qgso;
        gcggb = (cggb + cgdo + cgso + PP(cgbo)) * ag0;
        gcgdb = (cgsb - cgdo) * ag0;
        gcgsb = (cgdb - cgso) * ag0;
        gcgbb = -(gcggb + gcgdb + gcgsb);
        gcdgb = -(cggb + cbgb + cdgb + cgdo) * ag0;
        gcsgb = (cdgb - cgso) * ag0;
        gcbgb = (cbgb - PP(cgbo)) * ag0;
        gcdgmb = gcsgmb = gcbgmb = 0.0;
        qgb = PP(cgbo) * vgb;
        qgate += qgdo + qgso + qgb;
        qbulk -= qgb;
        qdrn = -(qgate + qbulk + qsrc);
        gcddb = (capbd + cgdo - (cgsb + cbsb + cdsb)) * ag0;
        gcdsb = -(cgdb + cbdb + cddb) * ag0;
        gcsdb = cdsb * ag0;
        gcssb = (cddb + capbs + cgso) * ag0;
        gcdbb = -(gcdgb + gcddb + gcdsb + gcdgmb);
        gcsbb = -(gcsgb + gcsdb + gcssb + gcsgmb);
        gcbdb = (cbsb - capbd) * ag0;
        gcbsb = (cbdb - capbs) * ag0;
        gcdbdb = 0.0;
        gcsbsb = 0.0;
        gcbbb = -(gcbgb + gcbdb + gcbsb + gcbgmb);
        ggtg = ggtd = ggtb = ggts = 0.0;
        sxpart = 0.4;
        dxpart = 0.6;
        ddxpart_dVd = ddxpart_dVg = ddxpart_dVb = ddxpart_dVs = 0.0;
        dsxpart_dVd = dsxpart_dVg = dsxpart_dVb = dsxpart_dVs = 0.0;
      }
    }
    {
      if (mode > 0) {
        CAPcggb = (cggb + cgdo + cgso + PP(cgbo));
        CAPcgdb = (cgdb - cgdo);
        CAPcgsb = (cgsb - cgso);
        CAPcgbb = -(CAPcggb + CAPcgdb + CAPcgsb);
        CAPcdgb = (cdgb - cgdo);
        CAPcsgb = -(cggb + cbgb + cdgb + cgso);
        CAPcbgb = (cbgb - PP(cgbo));
        CAPcdgmb = CAPcsgmb = CAPcbgmb = 0.0;
        CAPcddb = (cddb + capbd + cgdo);
        CAPcdsb = cdsb;
        CAPcsdb = -(cgdb + cbdb + cddb);
        CAPcssb = (capbs + cgso - (cgsb + cbsb + cdsb));
        CAPcdbb = -(CAPcdgb + CAPcddb + CAPcdsb + CAPcdgmb);
        CAPcsbb = -(CAPcsgb + CAPcsdb + CAPcssb + CAPcsgmb);
        CAPcbdb = (cbdb - capbd);
        CAPcbsb = (cbsb - capbs);
        CAPcdbdb = 0.0;
        CAPcsbsb = 0.0;
        CAPcbbb = -(CAPcbdb + CAPcbgb + CAPcbsb + CAPcbgmb);
      } else {
        CAPcggb = (cggb + cgdo + cgso + PP(cgbo));
        CAPcgdb = (cgsb - cgdo);
        CAPcgsb = (cgdb - cgso);
        CAPcgbb = -(CAPcggb + CAPcgdb + CAPcgsb);
        CAPcdgb = -(cggb + cbgb + cdgb + cgdo);
        CAPcsgb = (cdgb - cgso);
        CAPcbgb = (cbgb - PP(cgbo));
        CAPcdgmb = CAPcsgmb = CAPcbgmb = 0.0;
        CAPcddb = (capbd + cgdo - (cgsb + cbsb + cdsb));
        CAPcdsb = -(cgdb + cbdb + cddb);
        CAPcsdb = cdsb;
        CAPcssb = (cddb + capbs + cgso);
        CAPcdbb = -(CAPcdgb + CAPcddb + CAPcdsb + CAPcdgmb);
        CAPcsbb = -(CAPcsgb + CAPcsdb + CAPcssb + CAPcsgmb);
        CAPcbdb = (cbsb - capbd);
        CAPcbsb = (cbdb - capbs);
        CAPcdbdb = 0.0;
        CAPcsbsb = 0.0;
        CAPcbbb = -(CAPcbgb + CAPcbdb + CAPcbsb + CAPcbgmb);
      }
    }
  }
  if (X && rdsMod == 1) {
    Idrain = gdtot * Vddp;
    Isource = gstot * Vssp;
  } else {
    Idrain = IV(drainConductance) * Vddp;
    Isource = IV(sourceConductance) * Vssp;
  }
  /* body network currents: the reference tests the model's rbodyMod here */
  if (X && ((xf >> XB_RBODYM) & 1) != 0) {
    Idbb = XP(grbdb) * Vdbb;
    Idbbp = XP(grbpd) * Vdbbp;
    Isbb = XP(grbsb) * Vsbb;
    Isbbp = XP(grbps) * Vsbbp;
    Ibpb = XP(grbpb) * Vbpb;
  } else {
    Idbb = Idbbp = Isbb = Isbbp = Ibpb = 0.0;
  }

  // ---- Master::updateState: stored voltages and charges ----
  {
    double* sto = a.v.nextSto;
    const int32_t sl[12] = {li_store_vbd, li_store_vbs, li_store_vgs, li_store_vds,
                            li_store_vges, li_store_vgms, li_store_vdes, li_store_vses,
                            li_store_vdbs, li_store_vsbs, li_store_vdbd, li_store_von};
    const double sv[12] = {vbd, vbs, vgs, vds, vges, vgms, vdes, vses, vdbs, vsbs, vdbd, von};
    for (int k = 0; k < 12; ++k)
      if (sl[k] >= 0) sto[sl[k]] = sv[k];
  }
  qg = qgate;
  qd = qdrn - qbd;
  qb = (X && irbodyMod) ? qbulk : qbulk + qbd + qbs;
  {
    const int32_t lqb = stalid(a, A_qb, i), lqg = stalid(a, A_qg, i), lqd = stalid(a, A_qd, i);
    if (double* sta = a.v.nextSta) {
      if (lqg >= 0) sta[lqg] = qg;
      if (lqd >= 0) sta[lqd] = qd;
      if (lqb >= 0) sta[lqb] = qb;
    }
    if (!ss.dcopFlag && ss.initTranFlag && ss.newtonIter == 0) {
      if (double* cs = a.v.currSta) {
        if (lqg >= 0) cs[lqg] = qgate;
        if (lqd >= 0) cs[lqd] = qdrn - qbd;
        if (lqb >= 0) cs[lqb] = (X && irbodyMod) ? qbulk : qbulk + qbd + qbs;
      }
    }
    if (X) {
      // the option state slots follow qb, qg, qd in registration order
      int k = NSTA_BASE;
      const int32_t lqgmid = irgateMod == 3 ? stalid(a, k++, i) : -1;
      const int32_t lqbs = irbodyMod ? stalid(a, k++, i) : -1;
      const int32_t lqbd = irbodyMod ? stalid(a, k++, i) : -1;
      const int32_t lqcheq = itrnqsMod ? stalid(a, k++, i) : -1;
      const int32_t lqcdump = itrnqsMod ? stalid(a, k++, i) : -1;
      const bool first_tran = !ss.dcopFlag && ss.initTranFlag && ss.newtonIter == 0;
      for (int pass = 0; pass < 2; ++pass) {
        double* v = pass == 0 ? a.v.nextSta : (first_tran ? a.v.currSta : nullptr);
        if (!v) continue;
        if (lqgmid >= 0) v[lqgmid] = qgmid;
        if (lqbs >= 0) v[lqbs] = qbs;
        if (lqbd >= 0) v[lqbd] = qbd;
        if (lqcheq >= 0) v[lqcheq] = qcheq;
        if (lqcdump >= 0) v[lqcdump] = qdef * kScalingFactor;
      }
    }
  }
  st(a, S_VGB, i) = vgb;
  if (ChargeComputationNeeded) {
    st(a, S_QBS, i) = qbs;
    st(a, S_QBD, i) = qbd;
    st(a, S_CAPBS, i) = capbs;
    st(a, S_CAPBD, i) = capbd;
  }

  // ---- Master::loadDAEVectors ----
  Sink<MODE> s(a, i);
  const double Mp = IV(numberParallel);
  if constexpr (X) {
    // setupFVectorVars (:13525-13838)
    {

      const double dt = C.dtype;
      double tmp_gcrgd;
      T0 = 0.0;
      ceqdrn_Jdxp = 0.0; ceqbd_Jdxp = 0.0; ceqbs_Jdxp = 0.0;
      Istoteq_Jdxp = 0.0; Idtoteq_Jdxp = 0.0;
      Ibtoteq_Jdxp = 0.0; Igtoteq_Jdxp = 0.0;
      ceqgcrg_Jdxp = 0.0; ceqgstot_Jdxp = 0.0; ceqgdtot_Jdxp = 0.0;
      ceqjs_Jdxp = 0.0; ceqjd_Jdxp = 0.0;
      if (mode >= 0) {
        Gm = gm;
        Gmbs = gmbs;
        FwdSum = Gm + Gmbs;
        RevSum = 0.0;
        ceqdrn = dt * cdrain;
        ceqdrn_Jdxp = dt * (-gds * (vds - vds_orig) - Gm * (vgs - vgs_orig)
                               - Gmbs * (vbs - vbs_orig));
        ceqbd = dt * (csub + Igidl);
        ceqbd_Jdxp = dt * (-(gbds + ggidld) * (vds - vds_orig)
                              - (gbgs + ggidlg) * (vgs - vgs_orig)
                              - (gbbs + ggidlb) * (vbs - vbs_orig));
        ceqbs = dt * Igisl;
        ceqbs_Jdxp = dt * (+ggisls * (vds - vds_orig) - ggislg * (vgd - vgd_orig)
                              - ggislb * (vbd - vbd_orig));
        gbbdp = -(gbds);
        gbbsp = gbds + gbgs + gbbs;
        gbdpg = gbgs;
        gbdpdp = gbds;
        gbdpb = gbbs;
        gbdpsp = -(gbdpg + gbdpdp + gbdpb);
        gbspg = 0.0;
        gbspdp = 0.0;
        gbspb = 0.0;
        gbspsp = 0.0;
        if (igcMod) {
          gIstotg = gIgsg + gIgcsg;
          gIstotd = gIgcsd;
          gIstots = gIgss + gIgcss;
          gIstotb = gIgcsb;
          Istoteq = dt * (Igs + Igcs);
          Istoteq_Jdxp = dt * (-gIstotg * (vgs - vgs_orig) - gIgcsd * (vds - vds_orig)
                                  - gIgcsb * (vbs - vbs_orig));
          gIdtotg = gIgdg + gIgcdg;
          gIdtotd = gIgdd + gIgcdd;
          gIdtots = gIgcds;
          gIdtotb = gIgcdb;
          Idtoteq = dt * (Igd + Igcd);
          Idtoteq_Jdxp = dt * (-gIgdg * (vgd - vgd_orig) - gIgcdg * (vgs - vgs_orig)
                                  - gIgcdd * (vds - vds_orig) - gIgcdb * (vbs - vbs_orig));
        } else {
          gIstotg = gIstotd = gIstots = gIstotb = Istoteq = 0.0;
          gIdtotg = gIdtotd = gIdtots = gIdtotb = Idtoteq = 0.0;
          Istoteq_Jdxp = 0.0;
          Idtoteq_Jdxp = 0.0;
        }
        if (igbMod) {
          gIbtotg = gIgbg;
          gIbtotd = gIgbd;
          gIbtots = gIgbs;
          gIbtotb = gIgbb;
          Ibtoteq = dt * Igb;
          Ibtoteq_Jdxp = dt * (-gIgbg * (vgs - vgs_orig) - gIgbd * (vds - vds_orig)
                                  - gIgbb * (vbs - vbs_orig));
        } else {
          gIbtotg = gIbtotd = gIbtots = gIbtotb = Ibtoteq = 0.0;
          Ibtoteq_Jdxp = 0.0;
        }
        if ((igcMod != 0) || (igbMod != 0)) {
          gIgtotg = gIstotg + gIdtotg + gIbtotg;
          gIgtotd = gIstotd + gIdtotd + gIbtotd;
          gIgtots = gIstots + gIdtots + gIbtots;
          gIgtotb = gIstotb + gIdtotb + gIbtotb;
          Igtoteq = Istoteq + Idtoteq + Ibtoteq;
          Igtoteq_Jdxp = Istoteq_Jdxp + Idtoteq_Jdxp + Ibtoteq_Jdxp;
        } else {
          gIgtotg = gIgtotd = gIgtots = gIgtotb = Igtoteq = 0.0;
          Igtoteq_Jdxp = 0.0;
        }
        if (irgateMod == 2) T0 = vges - vgs;
        else if (irgateMod == 3) T0 = vgms - vgs;
        if (irgateMod > 1) {
          gcrgd = gcrgd * T0;
          gcrgg = gcrgg * T0;
          gcrgs = gcrgs * T0;
          gcrgb = gcrgb * T0;
          ceqgcrg = 0.0;
          ceqgcrg_Jdxp = -(gcrgd * (vds - vds_orig) + gcrgg * (vgs - vgs_orig)
                              + gcrgb * (vbs - vbs_orig));
          gcrgg -= gcrg;
        } else {
          ceqgcrg = gcrg = gcrgd = gcrgg = gcrgs = gcrgb = 0.0;
          ceqgcrg_Jdxp = 0.0;
        }
      } else {
        Gm = -gm;
        Gmbs = -gmbs;
        FwdSum = 0.0;
        RevSum = -(Gm + Gmbs);
        ceqdrn = -dt * cdrain;
        ceqdrn_Jdxp = -dt * (+gds * (vds - vds_orig) + Gm * (vgd - vgd_orig)
                                + Gmbs * (vbd - vbd_orig));
        ceqbs = dt * (csub + Igisl);
        ceqbs_Jdxp = dt * (+(gbds + ggisls) * (vds - vds_orig)
                              - (gbgs + ggislg) * (vgd - vgd_orig)
                              - (gbbs + ggislb) * (vbd - vbd_orig));
        ceqbd = dt * Igidl;
        ceqbd_Jdxp = dt * (-ggidld * (vds - vds_orig) - ggidlg * (vgs - vgs_orig)
                              - ggidlb * (vbs - vbs_orig));
        gbbsp = -(gbds);
        gbbdp = gbds + gbgs + gbbs;
        gbdpg = 0.0;
        gbdpsp = 0.0;
        gbdpb = 0.0;
        gbdpdp = 0.0;
        gbspg = gbgs;
        gbspsp = gbds;
        gbspb = gbbs;
        gbspdp = -(gbspg + gbspsp + gbspb);
        if (igcMod) {
          gIstotg = gIgsg + gIgcdg;
          gIstotd = gIgcds;
          gIstots = gIgss + gIgcdd;
          gIstotb = gIgcdb;
          Istoteq = dt * (Igs + Igcd);
          Istoteq_Jdxp = dt * (-gIgsg * (vgs - vgs_orig) - gIgcdg * (vgd - vgd_orig)
                                  + gIgcdd * (vds - vds_orig) - gIgcdb * (vbd - vbd_orig));
          gIdtotg = gIgdg + gIgcsg;
          gIdtotd = gIgdd + gIgcss;
          gIdtots = gIgcsd;
          gIdtotb = gIgcsb;
          Idtoteq = dt * (Igd + Igcs);
          Idtoteq_Jdxp = dt * (-(gIgdg + gIgcsg) * (vgd - vgd_orig)
                                  + gIgcsd * (vds - vds_orig) - gIgcsb * (vbd - vbd_orig));
        } else {
          gIstotg = gIstotd = gIstots = gIstotb = Istoteq = 0.0;
          gIdtotg = gIdtotd = gIdtots = gIdtotb = Idtoteq = 0.0;
          Istoteq_Jdxp = 0.0;
          Idtoteq_Jdxp = 0.0;
        }
        if (igbMod) {
          gIbtotg = gIgbg;
          gIbtotd = gIgbs;
          gIbtots = gIgbd;
          gIbtotb = gIgbb;
          Ibtoteq = dt * Igb;
          Ibtoteq_Jdxp = dt * (-gIgbg * (vgd - vgd_orig) + gIgbd * (vds - vds_orig)
                                  - gIgbb * (vbd - vbd_orig));
        } else {
          gIbtotg = gIbtotd = gIbtots = gIbtotb = Ibtoteq = 0.0;
          Ibtoteq_Jdxp = 0.0;
        }
        if ((igcMod != 0) || (igbMod != 0)) {
          gIgtotg = gIstotg + gIdtotg + gIbtotg;
          gIgtotd = gIstotd + gIdtotd + gIbtotd;
          gIgtots = gIstots + gIdtots + gIbtots;
          gIgtotb = gIstotb + gIdtotb + gIbtotb;
          Igtoteq = Istoteq + Idtoteq + Ibtoteq;
          Igtoteq_Jdxp = Istoteq_Jdxp + Idtoteq_Jdxp + Ibtoteq_Jdxp;
        } else {
          gIgtotg = gIgtotd = gIgtots = gIgtotb = Igtoteq = 0.0;
          Igtoteq_Jdxp = 0.0;
        }
        if (irgateMod == 2) T0 = vges - vgs;
        else if (irgateMod == 3) T0 = vgms - vgs;
        if (irgateMod > 1) {
          tmp_gcrgd = gcrgd;
          gcrgd = gcrgs * T0;
          gcrgg = gcrgg * T0;
          gcrgs = tmp_gcrgd * T0;
          gcrgb = gcrgb * T0;
          ceqgcrg = 0.0;
          ceqgcrg_Jdxp = -(gcrgg * (vgd - vgd_orig) - gcrgs * (vds - vds_orig)
                              + gcrgb * (vbd - vbd_orig));
          gcrgg -= gcrg;
        } else {
          /* the reverse branch leaves ceqgcrg_Jdxp at its zero */
          ceqgcrg = gcrg = gcrgd = gcrgg = gcrgs = gcrgb = 0.0;
        }
      }
      if (rdsMod == 1) {
        ceqgstot = 0.0;
        ceqgstot_Jdxp = dt * (gstotd * (vds - vds_orig) + gstotg * (vgs - vgs_orig)
                                 + gstotb * (vbs - vbs_orig));
        gstots = gstots - gstot;
        ceqgdtot = 0.0;
        ceqgdtot_Jdxp = -dt * (gdtotd * (vds - vds_orig) + gdtotg * (vgs - vgs_orig)
                                  + gdtotb * (vbs - vbs_orig));
        gdtotd = gdtotd - gdtot;
      } else {
        gstot = gstotd = gstotg = gstots = gstotb = ceqgstot = 0.0;
        gdtot = gdtotd = gdtotg = gdtots = gdtotb = ceqgdtot = 0.0;
        ceqgstot_Jdxp = 0.0;
        ceqgdtot_Jdxp = 0.0;
      }
      if (C.dtype > 0) {
        ceqjs = (cbs);
        ceqjs_Jdxp = (-gbs * (vbs_jct - vbs_jct_orig));
        ceqjd = (cbd);
        ceqjd_Jdxp = (-gbd * (vbd_jct - vbd_jct_orig));
      } else {
        ceqjs = -(cbs);
        ceqjs_Jdxp = (gbs * (vbs_jct - vbs_jct_orig));
        ceqjd = -(cbd);
        ceqjd_Jdxp = (gbd * (vbd_jct - vbd_jct_orig));
        ceqgcrg = -ceqgcrg;
        ceqgcrg_Jdxp = -ceqgcrg_Jdxp;
      }
    }
    const bool dc = ss.dcopFlag != 0;
    s.put(K_F, 0, -(ceqjd - ceqbd - ceqdrn + Idtoteq) * Mp);
    s.put(K_F, 1, -(-(-ceqgcrg + Igtoteq) * Mp));
    if (irgateMod == 1) {
      s.put(K_F, 2, -(-(Igate) * Mp));
      s.put(K_F, 4, -(Igate) * Mp);
    } else if (irgateMod == 2) {
      s.put(K_F, 2, -(-(Igate + ceqgcrg) * Mp));
      s.put(K_F, 4, -(Igate) * Mp);
    } else if (irgateMod == 3) {
      s.put(K_F, 2, -(-(Igate) * Mp));
      s.put(K_F, 3, -(-(IgateMid - Igate + ceqgcrg) * Mp));
      s.put(K_F, 4, -(IgateMid) * Mp);
    }
    if (!irbodyMod) {
      s.put(K_F, 6, -(ceqbd + ceqbs - ceqjd - ceqjs + Ibtoteq) * Mp);
      s.put(K_F, 9, -(ceqdrn - ceqbs + ceqjs + Istoteq) * Mp);
    } else {
      s.put(K_F, 5, -(-(ceqjd + Idbb + Idbbp) * Mp));
      s.put(K_F, 6, -(ceqbd + ceqbs + Ibtoteq + Idbbp + Isbbp - Ibpb) * Mp);
      s.put(K_F, 7, -(Isbb + Idbb + Ibpb) * Mp);
      s.put(K_F, 8, -(-(ceqjs + Isbb + Isbbp) * Mp));
      s.put(K_F, 9, -(ceqdrn - ceqbs + ceqjs + Istoteq) * Mp);
    }
    if (rdsMod) {
      s.put(K_F, 10, -(-ceqgdtot) * Mp);
      s.put(K_F, 11, -(ceqgstot) * Mp);
      s.put(K_F, 12, -(ceqgdtot) * Mp);
      s.put(K_F, 13, -(-ceqgstot) * Mp);
    }
    if (ifld(a, F_DEXIST, i)) {
      s.put(K_F, 14, -(-Idrain) * Mp);
      s.put(K_F, 15, -(Idrain) * Mp);
    }
    if (ifld(a, F_SEXIST, i)) {
      s.put(K_F, 16, -(-Isource) * Mp);
      s.put(K_F, 17, -(+Isource) * Mp);
    }
    // initial conditions: the branch equations in DC (:16382-16411); other
    // calls load nothing there (-0.0 leaves every sum unchanged)
    {
      const int32_t lb[3] = {node(a, N_IBS, i), node(a, N_IDS, i), node(a, N_IGS, i)};
      const int32_t lp[3] = {li_Body, li_Drain, li_GateExt};
      const double icv[3] = {XP(icVBS), XP(icVDS), XP(icVGS)};
      const int on[3] = {(xf >> XB_ICVBS) & 1, (xf >> XB_ICVDS) & 1, (xf >> XB_ICVGS) & 1};
      for (int k = 0; k < 3; ++k) {
        if (!on[k]) continue;
        const double coef = solv(a.v.nextSol, lb[k]);
        const double br = solv(a.v.nextSol, lp[k]) - solv(a.v.nextSol, li_Source) - icv[k];
        s.put(K_F, 18 + 3 * k, dc ? coef : -0.0);
        s.put(K_F, 19 + 3 * k, dc ? -coef : -0.0);
        s.put(K_F, 20 + 3 * k, dc ? br : -0.0);
      }
    }
    if (opt.voltageLimiterFlag) {
      const bool lim = !origFlag;
      const double z = -0.0;
      s.put(K_FLIM, 0, lim ? (ceqjd_Jdxp - ceqbd_Jdxp - ceqdrn_Jdxp + Idtoteq_Jdxp) * Mp : z);
      s.put(K_FLIM, 1, lim ? -((-ceqgcrg_Jdxp + Igtoteq_Jdxp) * Mp) : z);
      if (irgateMod == 1) {
        s.put(K_FLIM, 2, lim ? -((Igate_Jdxp) * Mp) : z);
        s.put(K_FLIM, 4, lim ? (Igate_Jdxp) * Mp : z);
      } else if (irgateMod == 2) {
        s.put(K_FLIM, 2, lim ? -((Igate_Jdxp + ceqgcrg_Jdxp) * Mp) : z);
        s.put(K_FLIM, 4, lim ? (Igate_Jdxp) * Mp : z);
      } else if (irgateMod == 3) {
        s.put(K_FLIM, 2, lim ? -((Igate_Jdxp) * Mp) : z);
        s.put(K_FLIM, 3, lim ? -((IgateMid_Jdxp - Igate_Jdxp + ceqgcrg_Jdxp) * Mp) : z);
        s.put(K_FLIM, 4, lim ? (IgateMid_Jdxp) * Mp : z);
      }
      if (!irbodyMod) {
        s.put(K_FLIM, 6, lim ? (ceqbd_Jdxp + ceqbs_Jdxp - ceqjd_Jdxp - ceqjs_Jdxp + Ibtoteq_Jdxp) * Mp : z);
        s.put(K_FLIM, 8, lim ? (ceqdrn_Jdxp - ceqbs_Jdxp + ceqjs_Jdxp + Istoteq_Jdxp) * Mp : z);
      } else {
        s.put(K_FLIM, 5, lim ? -((ceqjd_Jdxp) * Mp) : z);
        s.put(K_FLIM, 6, lim ? (ceqbd_Jdxp + ceqbs_Jdxp + Ibtoteq_Jdxp) * Mp : z);
        s.put(K_FLIM, 7, lim ? -((ceqjs_Jdxp) * Mp) : z);
        s.put(K_FLIM, 8, lim ? (ceqdrn_Jdxp - ceqbs_Jdxp + ceqjs_Jdxp + Istoteq_Jdxp) * Mp : z);
      }
      if (rdsMod) {
        s.put(K_FLIM, 9, lim ? -((ceqgdtot_Jdxp) * Mp) : z);
        s.put(K_FLIM, 10, lim ? (ceqgstot_Jdxp) * Mp : z);
        s.put(K_FLIM, 11, lim ? (ceqgdtot_Jdxp) * Mp : z);
        s.put(K_FLIM, 12, lim ? -((ceqgstot_Jdxp) * Mp) : z);
      }
    }
    // auxChargeCalculations (:12216-12347)
    {

      if (!ChargeComputationNeeded) {
        sxpart = (1.0 - (dxpart = (mode > 0) ? 0.4 : 0.6));
        ddxpart_dVd = ddxpart_dVg = ddxpart_dVb = ddxpart_dVs = 0.0;
        dsxpart_dVd = dsxpart_dVg = dsxpart_dVb = dsxpart_dVs = 0.0;
        Qeqqg_Jdxp = 0.0;
        Qeqqd_Jdxp = 0.0;
        Qeqqb_Jdxp = 0.0;
        Qeqqgmid_Jdxp = 0.0;
        Qeqqjs_Jdxp = 0.0;
        Qeqqjd_Jdxp = 0.0;
        if (itrnqsMod) {
          CoxWL = C.coxe * PP(weffCV) * IV(nf) * PP(leffCV);
          T1 = gcrg / CoxWL;
          gtau = T1 * kScalingFactor;
        } else {
          gtau = 0.0;
        }
      } else {
        Qeqqg_Jdxp = 0.0;
        if (!origFlag)
          Qeqqg_Jdxp = -CAPcggb * (vgb - vgb_orig) + CAPcgdb * (vbd - vbd_orig)
                          + CAPcgsb * (vbs - vbs_orig);
        Qeqqd_Jdxp = 0.0;
        if (!origFlag)
          Qeqqd_Jdxp = -CAPcdgb * (vgb - vgb_orig) - CAPcdgmb * (vgmb - vgmb_orig)
                          + (CAPcddb + CAPcdbdb) * (vbd - vbd_orig)
                          - CAPcdbdb * (vbd_jct - vbd_jct_orig) + CAPcdsb * (vbs - vbs_orig);
        Qeqqb_Jdxp = 0.0;
        if (!origFlag)
          Qeqqb_Jdxp = -CAPcbgb * (vgb - vgb_orig) - CAPcbgmb * (vgmb - vgmb_orig)
                          + CAPcbdb * (vbd - vbd_orig) + CAPcbsb * (vbs - vbs_orig);
        if (irgateMod == 3) {
          Qeqqgmid_Jdxp = 0.0;
          if (!origFlag)
            Qeqqgmid_Jdxp = +CAPcgmdb * (vbd - vbd_orig) + CAPcgmsb * (vbs - vbs_orig)
                               - CAPcgmgmb * (vgmb - vgmb_orig);
        } else {
          Qeqqgmid_Jdxp = 0.0;
        }
        if (irbodyMod) {
          Qeqqjs_Jdxp = 0.0;
          Qeqqjd_Jdxp = 0.0;
          if (!origFlag) {
            Qeqqjs_Jdxp = CAPcsbsb * (vbs_jct - vbs_jct_orig);
            Qeqqjd_Jdxp = CAPcdbdb * (vbd_jct - vbd_jct_orig);
          }
        }
        if (itrnqsMod) {
          T0 = ggtg * (vgb - vgb_orig) - ggtd * (vbd - vbd_orig)
               - ggts * (vbs - vbs_orig);
          Qeqqg_Jdxp += T0;
          T1 = qdef * gtau;
          Qeqqd_Jdxp -= dxpart * T0 + T1 * (ddxpart_dVg * (vgb - vgb_orig)
                                                  - ddxpart_dVd * (vbd - vbd_orig)
                                                  - ddxpart_dVs * (vbs - vbs_orig));
          Qqcheq_Jdxp = -(CAPcqgb * (vgb - vgb_orig) - CAPcqdb * (vbd - vbd_orig)
                             - CAPcqsb * (vbs - vbs_orig)) + T0;
        }
      }
    }
    double Qeqqg, Qeqqb, Qeqqd, Qeqqgmid = 0.0, Qeqqjs = 0.0, Qeqqjd = 0.0, Qqdef = 0.0, Qqcheq = 0.0;
    if (C.dtype > 0) {
      Qeqqg = qg;
      Qeqqd = qd;
      Qeqqb = qb;
      if (itrnqsMod) {
        Qqdef = qdef;
        Qqcheq = qcheq;
      }
      if (irbodyMod) {
        Qeqqjs = qbs;
        Qeqqjd = qbd;
      }
      if (irgateMod == 3) Qeqqgmid = qgmid;
    } else {
      Qeqqg = -qg;
      Qeqqd = -qd;
      Qeqqb = -qb;
      if (itrnqsMod) {
        Qqdef = -qdef;
        Qqcheq = -qcheq;
      }
      if (irbodyMod) {
        Qeqqjs = -qbs;
        Qeqqjd = -qbd;
      }
      if (irgateMod == 3) Qeqqgmid = -qgmid;
    }
    s.put(K_Q, 0, -(-Qeqqd) * Mp);
    s.put(K_Q, 1, -(-(Qeqqg) * Mp));
    if (irgateMod == 3) s.put(K_Q, 2, -(-(+Qeqqgmid) * Mp));
    if (!irbodyMod) {
      s.put(K_Q, 4, -(-Qeqqb) * Mp);
      s.put(K_Q, 6, -(+Qeqqg + Qeqqb + Qeqqd + Qeqqgmid) * Mp);
    } else {
      s.put(K_Q, 3, -(-(Qeqqjd) * Mp));
      s.put(K_Q, 4, -(-Qeqqb) * Mp);
      s.put(K_Q, 5, -(-(Qeqqjs) * Mp));
      s.put(K_Q, 6, -(Qeqqd + Qeqqg + Qeqqb + Qeqqjd + Qeqqjs + Qeqqgmid) * Mp);
    }
    if (itrnqsMod) s.put(K_Q, 7, -(Qqcheq - Qqdef) * Mp);
    if (opt.voltageLimiterFlag) {
      const bool lim = !origFlag;
      const double z = -0.0;
      s.put(K_QLIM, 0, lim ? (-Qeqqd_Jdxp) * Mp : z);
      s.put(K_QLIM, 1, lim ? -((Qeqqg_Jdxp) * Mp) : z);
      if (irgateMod == 3) s.put(K_QLIM, 2, lim ? -((+Qeqqgmid_Jdxp) * Mp) : z);
      if (!irbodyMod) {
        s.put(K_QLIM, 4, lim ? (-Qeqqb_Jdxp) * Mp : z);
        s.put(K_QLIM, 6, lim ? (+Qeqqg_Jdxp + Qeqqb_Jdxp + Qeqqd_Jdxp + Qeqqgmid_Jdxp) * Mp : z);
      } else {
        s.put(K_QLIM, 3, lim ? -((Qeqqjd_Jdxp) * Mp) : z);
        s.put(K_QLIM, 4, lim ? (-Qeqqb_Jdxp) * Mp : z);
        s.put(K_QLIM, 5, lim ? -((+Qeqqjs_Jdxp) * Mp) : z);
        s.put(K_QLIM, 6, lim ? (+Qeqqd_Jdxp + Qeqqg_Jdxp + Qeqqb_Jdxp + Qeqqjd_Jdxp + Qeqqjs_Jdxp
                                + Qeqqgmid_Jdxp) * Mp : z);
      }
      if (itrnqsMod) s.put(K_QLIM, 7, lim ? (Qqcheq_Jdxp) * Mp : z);
    }

    // ---- Master::loadDAEMatrices (option stamp lists) ----
    if (!irbodyMod) {
      gjbd = gbd;
      gjbs = gbs;
    } else {
      gjbd = gjbs = 0.0;
    }
    if (!rdsMod) {
      gdpr = IV(drainConductance);
      gspr = IV(sourceConductance);
    } else {
      gdpr = gspr = 0.0;
    }
    geltd = XP(grgeltd);
    T1 = qdef * gtau;
    if (irgateMod == 1) {
      s.put(K_DFDX, 0, (geltd) * Mp);
      s.put(K_DFDX, 1, -((geltd) * Mp));
      s.put(K_DFDX, 12, -((geltd) * Mp));
      s.put(K_DFDX, 14, (+geltd - ggtg + gIgtotg) * Mp);
      s.put(K_DFDX, 15, (-ggtd + gIgtotd) * Mp);
      s.put(K_DFDX, 16, (-ggts + gIgtots) * Mp);
      s.put(K_DFDX, 17, (-ggtb + gIgtotb) * Mp);
    } else if (irgateMod == 2) {
      s.put(K_DFDX, 0, (gcrg) * Mp);
      s.put(K_DFDX, 1, (gcrgg) * Mp);
      s.put(K_DFDX, 3, (gcrgd) * Mp);
      s.put(K_DFDX, 4, (gcrgs) * Mp);
      s.put(K_DFDX, 5, (gcrgb) * Mp);
      s.put(K_DFDX, 12, -((gcrg) * Mp));
      s.put(K_DFDX, 14, (-gcrgg - ggtg + gIgtotg) * Mp);
      s.put(K_DFDX, 15, (-gcrgd - ggtd + gIgtotd) * Mp);
      s.put(K_DFDX, 16, (-gcrgs - ggts + gIgtots) * Mp);
      s.put(K_DFDX, 17, (-gcrgb - ggtb + gIgtotb) * Mp);
    } else if (irgateMod == 3) {
      s.put(K_DFDX, 0, (geltd) * Mp);
      s.put(K_DFDX, 2, -((geltd) * Mp));
      s.put(K_DFDX, 6, -((geltd) * Mp));
      s.put(K_DFDX, 7, (geltd + gcrg) * Mp);
      s.put(K_DFDX, 8, (gcrgd) * Mp);
      s.put(K_DFDX, 9, (gcrgg) * Mp);
      s.put(K_DFDX, 10, (gcrgs) * Mp);
      s.put(K_DFDX, 11, (gcrgb) * Mp);
      s.put(K_DFDX, 13, -((gcrg) * Mp));
      s.put(K_DFDX, 14, (-gcrgg - ggtg + gIgtotg) * Mp);
      s.put(K_DFDX, 15, (-gcrgd - ggtd + gIgtotd) * Mp);
      s.put(K_DFDX, 16, (-gcrgs - ggts + gIgtots) * Mp);
      s.put(K_DFDX, 17, (-gcrgb - ggtb + gIgtotb) * Mp);
    } else {
      s.put(K_DFDX, 14, (-ggtg + gIgtotg) * Mp);
      s.put(K_DFDX, 15, (-ggtd + gIgtotd) * Mp);
      s.put(K_DFDX, 16, (-ggts + gIgtots) * Mp);
      s.put(K_DFDX, 17, (-ggtb + gIgtotb) * Mp);
    }
    if (rdsMod) {
      s.put(K_DFDX, 18, (gdtotg) * Mp);
      s.put(K_DFDX, 19, (gdtots) * Mp);
      s.put(K_DFDX, 20, (gdtotb) * Mp);
      s.put(K_DFDX, 21, (gstotd) * Mp);
      s.put(K_DFDX, 22, (gstotg) * Mp);
      s.put(K_DFDX, 23, (gstotb) * Mp);
    }
    s.put(K_DFDX, 24, (gdpr + gds + gbd + T1 * ddxpart_dVd - gdtotd + RevSum + gbdpdp + dxpart * ggtd - gIdtotd) * Mp);
    s.put(K_DFDX, 25, -((gdpr + gdtot) * Mp));
    s.put(K_DFDX, 26, (Gm - gdtotg + gbdpg - gIdtotg + dxpart * ggtg + T1 * ddxpart_dVg) * Mp);
    s.put(K_DFDX, 27, -((gds + gdtots - dxpart * ggts + gIdtots - T1 * ddxpart_dVs + FwdSum - gbdpsp) * Mp));
    s.put(K_DFDX, 28, -((gjbd + gdtotb - Gmbs - gbdpb + gIdtotb - T1 * ddxpart_dVb - dxpart * ggtb) * Mp));
    s.put(K_DFDX, 29, -((gdpr - gdtotd) * Mp));
    s.put(K_DFDX, 30, (gdpr + gdtot) * Mp);
    s.put(K_DFDX, 31, -((gds + gstotd + RevSum - gbspdp - T1 * dsxpart_dVd - sxpart * ggtd + gIstotd) * Mp));
    s.put(K_DFDX, 32, (-Gm - gstotg + gbspg + sxpart * ggtg + T1 * dsxpart_dVg - gIstotg) * Mp);
    s.put(K_DFDX, 33, (gspr + gds + gbs + T1 * dsxpart_dVs - gstots + FwdSum + gbspsp + sxpart * ggts - gIstots) * Mp);
    s.put(K_DFDX, 34, -((gspr + gstot) * Mp));
    s.put(K_DFDX, 35, -((gjbs + gstotb + Gmbs - gbspb - sxpart * ggtb - T1 * dsxpart_dVb + gIstotb) * Mp));
    s.put(K_DFDX, 36, -((gspr - gstots) * Mp));
    s.put(K_DFDX, 37, (gspr + gstot) * Mp);
    s.put(K_DFDX, 38, (-gjbd + gbbdp - gIbtotd) * Mp);
    s.put(K_DFDX, 39, (-gbgs - gIbtotg) * Mp);
    s.put(K_DFDX, 40, (-gjbs + gbbsp - gIbtots) * Mp);
    s.put(K_DFDX, 41, (gjbd + gjbs - gbbs - gIbtotb) * Mp);
    s.put(K_DFDX, 42, (ggidld) * Mp);
    s.put(K_DFDX, 43, (ggidlg) * Mp);
    s.put(K_DFDX, 44, -(((ggidlg + ggidld + ggidlb)) * Mp));
    s.put(K_DFDX, 45, (ggidlb) * Mp);
    s.put(K_DFDX, 46, -((ggidld) * Mp));
    s.put(K_DFDX, 47, -((ggidlg) * Mp));
    s.put(K_DFDX, 48, ((ggidlg + ggidld + ggidlb)) * Mp);
    s.put(K_DFDX, 49, -((ggidlb) * Mp));
    s.put(K_DFDX, 50, -(((ggisls + ggislg + ggislb)) * Mp));
    s.put(K_DFDX, 51, (ggislg) * Mp);
    s.put(K_DFDX, 52, (ggisls) * Mp);
    s.put(K_DFDX, 53, (ggislb) * Mp);
    s.put(K_DFDX, 54, ((ggislg + ggisls + ggislb)) * Mp);
    s.put(K_DFDX, 55, -((ggislg) * Mp));
    s.put(K_DFDX, 56, -((ggisls) * Mp));
    s.put(K_DFDX, 57, -((ggislb) * Mp));
    if (irbodyMod) {
      s.put(K_DFDX, 58, (-gbd) * Mp);
      s.put(K_DFDX, 59, -((gbs) * Mp));
      s.put(K_DFDX, 60, (-gbd) * Mp);
      s.put(K_DFDX, 61, (gbd + XP(grbpd) + XP(grbdb)) * Mp);
      s.put(K_DFDX, 62, -((XP(grbpd)) * Mp));
      s.put(K_DFDX, 63, -((XP(grbdb)) * Mp));
      s.put(K_DFDX, 64, -((XP(grbpd)) * Mp));
      s.put(K_DFDX, 65, -((XP(grbpb)) * Mp));
      s.put(K_DFDX, 66, -((XP(grbps)) * Mp));
      s.put(K_DFDX, 67, (XP(grbpd) + XP(grbps) + XP(grbpb)) * Mp);
      s.put(K_DFDX, 68, (-gbs) * Mp);
      s.put(K_DFDX, 69, -((XP(grbps)) * Mp));
      s.put(K_DFDX, 70, -((XP(grbsb)) * Mp));
      s.put(K_DFDX, 71, (gbs + XP(grbps) + XP(grbsb)) * Mp);
      s.put(K_DFDX, 72, -((XP(grbdb)) * Mp));
      s.put(K_DFDX, 73, -((XP(grbpb)) * Mp));
      s.put(K_DFDX, 74, -((XP(grbsb)) * Mp));
      s.put(K_DFDX, 75, (XP(grbsb) + XP(grbdb) + XP(grbpb)) * Mp);
    }
    if (itrnqsMod) {
      s.put(K_DFDX, 76, (gqdef + gtau) * Mp);
      s.put(K_DFDX, 77, (ggtg) * Mp);
      s.put(K_DFDX, 78, (ggtd) * Mp);
      s.put(K_DFDX, 79, (ggts) * Mp);
      s.put(K_DFDX, 80, (ggtb) * Mp);
      s.put(K_DFDX, 81, (dxpart * gtau) * Mp);
      s.put(K_DFDX, 82, (sxpart * gtau) * Mp);
      s.put(K_DFDX, 83, -((gtau) * Mp));
    }
    // initial conditions (:16840-16887): the DC form adds the branch
    // incidences, the transient form puts 1 on the branch diagonal (which
    // only this instance addresses: assigning it equals adding to zero)
    {
      const int on[3] = {(xf >> XB_ICVBS) & 1, (xf >> XB_ICVDS) & 1, (xf >> XB_ICVGS) & 1};
      for (int k = 0; k < 3; ++k) {
        if (!on[k]) continue;
        const int e = 84 + 5 * k;
        s.put(K_DFDX, e, dc ? 1.0 : -0.0);
        s.put(K_DFDX, e + 1, dc ? -1.0 : -0.0);
        s.put(K_DFDX, e + 2, dc ? 1.0 : -0.0);
        s.put(K_DFDX, e + 3, dc ? -1.0 : -0.0);
        s.put(K_DFDX, e + 4, dc ? -0.0 : 1.0);
      }
    }
    if (irgateMod == 3) {
      s.put(K_DQDX, 0, (+CAPcgmgmb) * Mp);
      s.put(K_DQDX, 1, (CAPcgmdb) * Mp);
      s.put(K_DQDX, 2, (CAPcgmsb) * Mp);
      s.put(K_DQDX, 3, (CAPcgmbb) * Mp);
      s.put(K_DQDX, 4, (CAPcdgmb) * Mp);
      s.put(K_DQDX, 5, (CAPcsgmb) * Mp);
      s.put(K_DQDX, 6, (CAPcbgmb) * Mp);
    }
    s.put(K_DQDX, 7, (CAPcggb) * Mp);
    s.put(K_DQDX, 8, (CAPcgdb) * Mp);
    s.put(K_DQDX, 9, (CAPcgsb) * Mp);
    s.put(K_DQDX, 10, (CAPcgbb) * Mp);
    s.put(K_DQDX, 11, (CAPcddb) * Mp);
    s.put(K_DQDX, 12, (+CAPcdgb) * Mp);
    s.put(K_DQDX, 13, -((-CAPcdsb) * Mp));
    s.put(K_DQDX, 14, -((-CAPcdbb) * Mp));
    s.put(K_DQDX, 15, -((-CAPcsdb) * Mp));
    s.put(K_DQDX, 16, (CAPcsgb) * Mp);
    s.put(K_DQDX, 17, (CAPcssb) * Mp);
    s.put(K_DQDX, 18, -((-CAPcsbb) * Mp));
    s.put(K_DQDX, 19, (CAPcbdb) * Mp);
    s.put(K_DQDX, 20, (CAPcbgb) * Mp);
    s.put(K_DQDX, 21, (CAPcbsb) * Mp);
    s.put(K_DQDX, 22, (CAPcbbb) * Mp);
    if (irbodyMod) {
      s.put(K_DQDX, 23, (CAPcdbdb) * Mp);
      s.put(K_DQDX, 24, -((-CAPcsbsb) * Mp));
      s.put(K_DQDX, 25, (CAPcdbdb) * Mp);
      s.put(K_DQDX, 26, (-CAPcdbdb) * Mp);
      s.put(K_DQDX, 27, (CAPcsbsb) * Mp);
      s.put(K_DQDX, 28, (-CAPcsbsb) * Mp);
    }
    if (itrnqsMod) {
      s.put(K_DQDX, 29, (-CAPcqgb) * Mp);
      s.put(K_DQDX, 30, (-CAPcqdb) * Mp);
      s.put(K_DQDX, 31, (-CAPcqsb) * Mp);
      s.put(K_DQDX, 32, (-CAPcqbb) * Mp);
    }
    return;
  }
  {
    const double dt = C.dtype;
    ceqdrn_Jdxp = 0.0; ceqbd_Jdxp = 0.0; ceqbs_Jdxp = 0.0;
    Istoteq_Jdxp = 0.0; Idtoteq_Jdxp = 0.0;
    Ibtoteq_Jdxp = 0.0; Igtoteq_Jdxp = 0.0;
    ceqgcrg_Jdxp = 0.0; ceqgstot_Jdxp = 0.0; ceqgdtot_Jdxp = 0.0;
    ceqjs_Jdxp = 0.0; ceqjd_Jdxp = 0.0;
    if (mode >= 0) {
      Gm = gm;
      Gmbs = gmbs;
      FwdSum = Gm + Gmbs;
      RevSum = 0.0;
      ceqdrn = dt * cdrain;
      ceqdrn_Jdxp = dt * (-gds * (vds - vds_orig) - Gm * (vgs - vgs_orig)
                             - Gmbs * (vbs - vbs_orig));
      ceqbd = dt * (csub + Igidl);
      ceqbd_Jdxp = dt * (-(gbds + ggidld) * (vds - vds_orig)
                            - (gbgs + ggidlg) * (vgs - vgs_orig)
                            - (gbbs + ggidlb) * (vbs - vbs_orig));
      ceqbs = dt * Igisl;
      ceqbs_Jdxp = dt * (+ggisls * (vds - vds_orig) - ggislg * (vgd - vgd_orig)
                            - ggislb * (vbd - vbd_orig));
      gbbdp = -(gbds);
      gbbsp = gbds + gbgs + gbbs;
      gbdpg = gbgs;
      gbdpdp = gbds;
      gbdpb = gbbs;
      gbdpsp = -(gbdpg + gbdpdp + gbdpb);
      gbspg = 0.0;
      gbspdp = 0.0;
      gbspb = 0.0;
      gbspsp = 0.0;
    } else {
      Gm = -gm;
      Gmbs = -gmbs;
      FwdSum = 0.0;
      RevSum = -(Gm + Gmbs);
      ceqdrn = -dt * cdrain;
      ceqdrn_Jdxp = -dt * (+gds * (vds - vds_orig) + Gm * (vgd - vgd_orig)
                              + Gmbs * (vbd - vbd_orig));
      ceqbs = dt * (csub + Igisl);
      ceqbs_Jdxp = dt * (+(gbds + ggisls) * (vds - vds_orig)
                            - (gbgs + ggislg) * (vgd - vgd_orig)
                            - (gbbs + ggislb) * (vbd - vbd_orig));
      ceqbd = dt * Igidl;
      ceqbd_Jdxp = dt * (-ggidld * (vds - vds_orig) - ggidlg * (vgs - vgs_orig)
                            - ggidlb * (vbs - vbs_orig));
      gbbsp = -(gbds);
      gbbdp = gbds + gbgs + gbbs;
      gbdpg = 0.0;
      gbdpsp = 0.0;
      gbdpb = 0.0;
      gbdpdp = 0.0;
      gbspg = gbgs;
      gbspsp = gbds;
      gbspb = gbbs;
      gbspdp = -(gbspg + gbspsp + gbspb);
    }
    gIstotg = gIstotd = gIstots = gIstotb = Istoteq = 0.0;
    gIdtotg = gIdtotd = gIdtots = gIdtotb = Idtoteq = 0.0;
    gIbtotg = gIbtotd = gIbtots = gIbtotb = Ibtoteq = 0.0;
    gIgtotg = gIgtotd = gIgtots = gIgtotb = Igtoteq = 0.0;
    ceqgcrg = gcrg = gcrgd = gcrgg = gcrgs = gcrgb = 0.0;
    gstot = gstotd = gstotg = gstots = gstotb = ceqgstot = 0.0;
    gdtot = gdtotd = gdtotg = gdtots = gdtotb = ceqgdtot = 0.0;
    if (C.dtype > 0) {
      ceqjs = (cbs);
      ceqjs_Jdxp = (-gbs * (vbs_jct - vbs_jct_orig));
      ceqjd = (cbd);
      ceqjd_Jdxp = (-gbd * (vbd_jct - vbd_jct_orig));
    } else {
      ceqjs = -(cbs);
      ceqjs_Jdxp = (gbs * (vbs_jct - vbs_jct_orig));
      ceqjd = -(cbd);
      ceqjd_Jdxp = (gbd * (vbd_jct - vbd_jct_orig));
      ceqgcrg = -ceqgcrg;
      ceqgcrg_Jdxp = -ceqgcrg_Jdxp;
    }
  }
  s.put(K_F, 0, -(ceqjd - ceqbd - ceqdrn + Idtoteq) * Mp);
  s.put(K_F, 1, -(-(-ceqgcrg + Igtoteq) * Mp));
  s.put(K_F, 2, -(ceqbd + ceqbs - ceqjd - ceqjs + Ibtoteq) * Mp);
  s.put(K_F, 3, -(ceqdrn - ceqbs + ceqjs + Istoteq) * Mp);
  if (ifld(a, F_DEXIST, i)) {
    s.put(K_F, 4, -(-Idrain) * Mp);
    s.put(K_F, 5, -(Idrain) * Mp);
  }
  if (ifld(a, F_SEXIST, i)) {
    s.put(K_F, 6, -(-Isource) * Mp);
    s.put(K_F, 7, -(+Isource) * Mp);
  }
  // Limiter corrections are loaded only when limiting moved a voltage; an
  // instance that did not limit adds -0.0, which leaves every sum unchanged.
  if (opt.voltageLimiterFlag) {
    const bool lim = !origFlag;
    s.put(K_FLIM, 0, lim ? (ceqjd_Jdxp - ceqbd_Jdxp - ceqdrn_Jdxp + Idtoteq_Jdxp) * Mp : -0.0);
    s.put(K_FLIM, 1, lim ? -((-ceqgcrg_Jdxp + Igtoteq_Jdxp) * Mp) : -0.0);
    s.put(K_FLIM, 2, lim ? (ceqbd_Jdxp + ceqbs_Jdxp - ceqjd_Jdxp - ceqjs_Jdxp + Ibtoteq_Jdxp) * Mp : -0.0);
    s.put(K_FLIM, 3, lim ? (ceqdrn_Jdxp - ceqbs_Jdxp + ceqjs_Jdxp + Istoteq_Jdxp) * Mp : -0.0);
  }
  {
    if (!ChargeComputationNeeded) {
      sxpart = (1.0 - (dxpart = (mode > 0) ? 0.4 : 0.6));
      ddxpart_dVd = ddxpart_dVg = ddxpart_dVb = ddxpart_dVs = 0.0;
      dsxpart_dVd = dsxpart_dVg = dsxpart_dVb = dsxpart_dVs = 0.0;
      Qeqqg_Jdxp = 0.0;
      Qeqqd_Jdxp = 0.0;
      Qeqqb_Jdxp = 0.0;
      Qeqqgmid_Jdxp = 0.0;
      Qeqqjs_Jdxp = 0.0;
      Qeqqjd_Jdxp = 0.0;
      gtau = 0.0;
    } else {
      Qeqqg_Jdxp = 0.0;
      if (!origFlag)
        Qeqqg_Jdxp = -CAPcggb * (vgb - vgb_orig) + CAPcgdb * (vbd - vbd_orig)
                        + CAPcgsb * (vbs - vbs_orig);
      Qeqqd_Jdxp = 0.0;
      if (!origFlag)
        Qeqqd_Jdxp = -CAPcdgb * (vgb - vgb_orig) - CAPcdgmb * (vgmb - vgmb_orig)
                        + (CAPcddb + CAPcdbdb) * (vbd - vbd_orig)
                        - CAPcdbdb * (vbd_jct - vbd_jct_orig) + CAPcdsb * (vbs - vbs_orig);
      Qeqqb_Jdxp = 0.0;
      if (!origFlag)
        Qeqqb_Jdxp = -CAPcbgb * (vgb - vgb_orig) - CAPcbgmb * (vgmb - vgmb_orig)
                        + CAPcbdb * (vbd - vbd_orig) + CAPcbsb * (vbs - vbs_orig);
      Qeqqgmid_Jdxp = 0.0;
    }
  }
  double Qeqqg, Qeqqb, Qeqqd;
  const double Qeqqgmid = 0.0;
  if (C.dtype > 0) {
    Qeqqg = qg;
    Qeqqd = qd;
    Qeqqb = qb;
  } else {
    Qeqqg = -qg;
    Qeqqd = -qd;
    Qeqqb = -qb;
  }
  s.put(K_Q, 0, -(-Qeqqd) * Mp);
  s.put(K_Q, 1, -(-(Qeqqg) * Mp));
  s.put(K_Q, 2, -(-Qeqqb) * Mp);
  s.put(K_Q, 3, -(+Qeqqg + Qeqqb + Qeqqd + Qeqqgmid) * Mp);
  if (opt.voltageLimiterFlag) {
    const bool lim = !origFlag;
    s.put(K_QLIM, 0, lim ? (-Qeqqd_Jdxp) * Mp : -0.0);
    s.put(K_QLIM, 1, lim ? -((Qeqqg_Jdxp) * Mp) : -0.0);
    s.put(K_QLIM, 2, lim ? (-Qeqqb_Jdxp) * Mp : -0.0);
    s.put(K_QLIM, 3, lim ? (+Qeqqg_Jdxp + Qeqqb_Jdxp + Qeqqd_Jdxp + Qeqqgmid_Jdxp) * Mp : -0.0);
  }

  // ---- Master::loadDAEMatrices ----
  gjbd = gbd;
  gjbs = gbs;
  gdpr = IV(drainConductance);
  gspr = IV(sourceConductance);
  T1 = qdef * gtau;
    s.put(K_DFDX, 0, (-ggtg + gIgtotg) * Mp);
    s.put(K_DFDX, 1, (-ggtd + gIgtotd) * Mp);
    s.put(K_DFDX, 2, (-ggts + gIgtots) * Mp);
    s.put(K_DFDX, 3, (-ggtb + gIgtotb) * Mp);
    s.put(K_DFDX, 4, (gdpr + gds + gbd + T1 * ddxpart_dVd - gdtotd + RevSum + gbdpdp + dxpart * ggtd - gIdtotd) * Mp);
    s.put(K_DFDX, 5, -((gdpr + gdtot) * Mp));
    s.put(K_DFDX, 6, (Gm - gdtotg + gbdpg - gIdtotg + dxpart * ggtg + T1 * ddxpart_dVg) * Mp);
    s.put(K_DFDX, 7, -((gds + gdtots - dxpart * ggts + gIdtots - T1 * ddxpart_dVs + FwdSum - gbdpsp) * Mp));
    s.put(K_DFDX, 8, -((gjbd + gdtotb - Gmbs - gbdpb + gIdtotb - T1 * ddxpart_dVb - dxpart * ggtb) * Mp));
    s.put(K_DFDX, 9, -((gdpr - gdtotd) * Mp));
    s.put(K_DFDX, 10, (gdpr + gdtot) * Mp);
    s.put(K_DFDX, 11, -((gds + gstotd + RevSum - gbspdp - T1 * dsxpart_dVd - sxpart * ggtd + gIstotd) * Mp));
    s.put(K_DFDX, 12, (-Gm - gstotg + gbspg + sxpart * ggtg + T1 * dsxpart_dVg - gIstotg) * Mp);
    s.put(K_DFDX, 13, (gspr + gds + gbs + T1 * dsxpart_dVs - gstots + FwdSum + gbspsp + sxpart * ggts - gIstots) * Mp);
    s.put(K_DFDX, 14, -((gspr + gstot) * Mp));
    s.put(K_DFDX, 15, -((gjbs + gstotb + Gmbs - gbspb - sxpart * ggtb - T1 * dsxpart_dVb + gIstotb) * Mp));
    s.put(K_DFDX, 16, -((gspr - gstots) * Mp));
    s.put(K_DFDX, 17, (gspr + gstot) * Mp);
    s.put(K_DFDX, 18, (-gjbd + gbbdp - gIbtotd) * Mp);
    s.put(K_DFDX, 19, (-gbgs - gIbtotg) * Mp);
    s.put(K_DFDX, 20, (-gjbs + gbbsp - gIbtots) * Mp);
    s.put(K_DFDX, 21, (gjbd + gjbs - gbbs - gIbtotb) * Mp);
    s.put(K_DFDX, 22, (ggidld) * Mp);
    s.put(K_DFDX, 23, (ggidlg) * Mp);
    s.put(K_DFDX, 24, -(((ggidlg + ggidld + ggidlb)) * Mp));
    s.put(K_DFDX, 25, (ggidlb) * Mp);
    s.put(K_DFDX, 26, -((ggidld) * Mp));
    s.put(K_DFDX, 27, -((ggidlg) * Mp));
    s.put(K_DFDX, 28, ((ggidlg + ggidld + ggidlb)) * Mp);
    s.put(K_DFDX, 29, -((ggidlb) * Mp));
    s.put(K_DFDX, 30, -(((ggisls + ggislg + ggislb)) * Mp));
    s.put(K_DFDX, 31, (ggislg) * Mp);
    s.put(K_DFDX, 32, (ggisls) * Mp);
    s.put(K_DFDX, 33, (ggislb) * Mp);
    s.put(K_DFDX, 34, ((ggislg + ggisls + ggislb)) * Mp);
    s.put(K_DFDX, 35, -((ggislg) * Mp));
    s.put(K_DFDX, 36, -((ggisls) * Mp));
    s.put(K_DFDX, 37, -((ggislb) * Mp));
    s.put(K_DQDX, 0, (CAPcggb) * Mp);
    s.put(K_DQDX, 1, (CAPcgdb) * Mp);
    s.put(K_DQDX, 2, (CAPcgsb) * Mp);
    s.put(K_DQDX, 3, (CAPcgbb) * Mp);
    s.put(K_DQDX, 4, (CAPcddb) * Mp);
    s.put(K_DQDX, 5, (+CAPcdgb) * Mp);
    s.put(K_DQDX, 6, -((-CAPcdsb) * Mp));
    s.put(K_DQDX, 7, -((-CAPcdbb) * Mp));
    s.put(K_DQDX, 8, -((-CAPcsdb) * Mp));
    s.put(K_DQDX, 9, (CAPcsgb) * Mp);
    s.put(K_DQDX, 10, (CAPcssb) * Mp);
    s.put(K_DQDX, 11, -((-CAPcsbb) * Mp));
    s.put(K_DQDX, 12, (CAPcbdb) * Mp);
    s.put(K_DQDX, 13, (CAPcbgb) * Mp);
    s.put(K_DQDX, 14, (CAPcbsb) * Mp);
    s.put(K_DQDX, 15, (CAPcbbb) * Mp);
}

// The stored-mode kernel is held to two waves per SIMD: the eval's peak
// (about 135 live doubles in the DC model) then spills to scratch, and the
// second wave hides more memory latency than the spills cost (update_state 7%
// faster on 2^20 instances).  The fused kernel serves small latency-bound
// launches and keeps every value in registers.
//
// The stored kernel with LDSCARD first stages the cards of its block's table
// (the distinct models of its instances, at most kBlockCards) in LDS, so the
// eval's card reads are LDS reads instead of L1/L2 hits.
#ifndef XA_B4_STORED_WAVES
#define XA_B4_STORED_WAVES 2   // waves per SIMD of the stored-mode update (a tuning knob)
#endif
template <bool FUSED, bool LDSCARD, int CM>
__global__ void __launch_bounds__(256, FUSED ? 1 : XA_B4_STORED_WAVES) k_update(KArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (LDSCARD) {
    constexpr int kStride = kCardStride;
    __shared__ double s_card[kBlockCards * kStride];
    if (FUSED) {   // every card of the group (at most kBlockCards)
      for (int k = threadIdx.x; k < a.n_models * kStride; k += blockDim.x) s_card[k] = a.mcard[k];
    } else {
      const int32_t* tab = a.blk_cards + (int64_t)blockIdx.x * (kBlockCards + 1);
      const int used = tab[0];   // 1 <= used <= kBlockCards
      for (int k = threadIdx.x; k < used * kStride; k += blockDim.x)
        s_card[k] = a.mcard[(int64_t)tab[1 + k / kStride] * kStride + k % kStride];
    }
    __syncthreads();
    eval<FUSED ? 1 : 0, true, CM, !FUSED>(a, i, s_card);
  } else {
    eval<FUSED ? 1 : 0, false, CM, !FUSED>(a, i, nullptr);
  }
  if (FUSED) fused_writeout(a);
}

// Persistent form of the stored-mode update (xa_group_set_persistent): a
// resident grid whose workgroups loop over the instance tiles, tile t on
// workgroup t mod gridDim.x, staging each tile's card table in LDS.  The
// eval per instance is the stored kernel's, so the bits are the same.
template <bool LDSCARD, int CM>
__global__ void __launch_bounds__(256, XA_B4_STORED_WAVES) k_update_persist(KArgs a) {
  constexpr int kStride = kCardStride;
  __shared__ double s_card[LDSCARD ? kBlockCards * kStride : 1];
  const int64_t tiles = (a.n + blockDim.x - 1) / blockDim.x;
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int64_t i = t * blockDim.x + threadIdx.x;
    if (LDSCARD) {
      const int32_t* tab = a.blk_cards + t * (kBlockCards + 1);
      const int used = tab[0];   // 1 <= used <= kBlockCards
      for (int k = threadIdx.x; k < used * kStride; k += blockDim.x)
        s_card[k] = a.mcard[(int64_t)tab[1 + k / kStride] * kStride + k % kStride];
      __syncthreads();
      eval<0, true, CM, true>(a, i, s_card);
      __syncthreads();   // every lane done with this tile's cards
    } else {
      eval<0, false, CM, true>(a, i, nullptr);
    }
  }
}

// Small launches (at most about one wavefront per SIMD over the GPU): the
// stored-mode update with the block's cards in LDS, compiled for one wave per
// SIMD so that every value stays in registers -- no spills, and no section
// reloads of the parameter columns.  Such a launch has no second wave to hide
// a latency behind, so each scratch or column reload would sit on its
// critical path.
constexpr int64_t kSmallUpdate = 65536;   // instances
template <int CM>
__global__ void __launch_bounds__(256, 1) k_update_small(KArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int kStride = kCardStride;
  __shared__ double s_card[kBlockCards * kStride];
  const int32_t* tab = a.blk_cards + (int64_t)blockIdx.x * (kBlockCards + 1);
  const int used = tab[0];   // 1 <= used <= kBlockCards
  for (int k = threadIdx.x; k < used * kStride; k += blockDim.x)
    s_card[k] = a.mcard[(int64_t)tab[1 + k / kStride] * kStride + k % kStride];
  __syncthreads();
  eval<0, true, CM, false>(a, i, s_card);
}

// The option kernel: the stored-mode update of option groups (every option
// branch compiled in), one wave per SIMD so that its larger register set
// stays in registers; the block's cards in LDS when its card tables exist.
template <bool LDSCARD>
__global__ void __launch_bounds__(256, 1) k_update_x(KArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (LDSCARD) {
    constexpr int kStride = kCardStride;
    __shared__ double s_card[kBlockCards * kStride];
    const int32_t* tab = a.blk_cards + (int64_t)blockIdx.x * (kBlockCards + 1);
    const int used = tab[0];   // 1 <= used <= kBlockCards
    for (int k = threadIdx.x; k < used * kStride; k += blockDim.x)
      s_card[k] = a.mcard[(int64_t)tab[1 + k / kStride] * kStride + k % kStride];
    __syncthreads();
    eval<0, true, 0, false, true>(a, i, s_card);
  } else {
    eval<0, false, 0, false, true>(a, i, nullptr);
  }
}

// Pattern-fused load (every kind a pattern plan): the block stages its output
// spans in LDS, every instance adds its stamps into its own span in stamp
// order as the eval produces them, and the block writes the spans back.  No
// value buffers and no scatter kernels; one wave per SIMD (the spans and the
// eval's registers).
template <bool LDSCARD, int CM>
__global__ void __launch_bounds__(256, 1) k_update_pattern(KArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int kStride = kCardStride;
  __shared__ double s_card[LDSCARD ? kBlockCards * kStride : 1];
  pattern_stage(a, true);
  if (LDSCARD) {
    const int32_t* tab = a.blk_cards + (int64_t)blockIdx.x * (kBlockCards + 1);
    const int used = tab[0];
    for (int k = threadIdx.x; k < used * kStride; k += blockDim.x)
      s_card[k] = a.mcard[(int64_t)tab[1 + k / kStride] * kStride + k % kStride];
  }
  __syncthreads();
  eval<2, LDSCARD, CM>(a, i, s_card);
  __syncthreads();
  pattern_stage(a, false);
}

#undef PP
#undef IV

// The largest card mode whose columns are all card-uniform in this group.
static int card_mode(const Group& g) {
  for (int cm = 2; cm > 0; --cm) {
    bool ok = true;
    for (int f = 0; f < ND && ok; ++f)
      ok = !card_col(cm, f) || ((g.ucol[f >> 6] >> (f & 63)) & 1);
    if (ok) return cm;
  }
  return 0;
}

static int card_columns(const Group& g) {
  const int cm = card_mode(g);
  int cols = 0;
  for (int f = 0; f < ND; ++f) cols += card_col(cm, f) ? 1 : 0;
  return cols;
}

static int launch_update(Group& g, const xa_solver_state& ss, const xa_vectors& v, bool fused) {
  if (!v.nextSol || !v.nextSto) return fail(XA_ERR_ARG, "bsim4 update: nextSol/nextSto required");
  if (ss.newtonIter == 0 && (!ss.dcopFlag || (ss.locaEnabledFlag && ss.dcopFlag)) && !v.currSto)
    return fail(XA_ERR_ARG, "bsim4 update: currSto required");
  if (ss.inputOPFlag && ss.initJctFlag && !v.flagSol)
    return fail(XA_ERR_ARG, "bsim4 update: flagSol required with inputOPFlag");
  KArgs a;
  fill_kargs(g, ss, v, fused, a);
  static_assert(sizeof(DevCard) % sizeof(double) == 0, "cards are whole doubles");
  // the tables exist for the current update block when no block uses more
  // than kBlockCards models (built at finalize and by the autotuner)
  const bool lds_cards = g.card_tables && g.card_block == g.split_block &&
                         a.mstride == kCardStride;
  const int cm = card_mode(g);
  const void* stored[2][3] = {
      {(const void*)k_update<false, false, 0>, (const void*)k_update<false, false, 1>,
       (const void*)k_update<false, false, 2>},
      {(const void*)k_update<false, true, 0>, (const void*)k_update<false, true, 1>,
       (const void*)k_update<false, true, 2>}};
  const void* fused_k[2][3] = {
      {(const void*)k_update<true, false, 0>, (const void*)k_update<true, false, 1>,
       (const void*)k_update<true, false, 2>},
      {(const void*)k_update<true, true, 0>, (const void*)k_update<true, true, 1>,
       (const void*)k_update<true, true, 2>}};
  // the fused kernel stages the cards in LDS beside its stamp rows when both fit
  const bool fused_lds = a.n_models <= kBlockCards && a.mstride == kCardStride &&
                         g.fused_lds + sizeof(double) * kBlockCards * kCardStride <= kMaxLds;
  const void* small[3] = {(const void*)k_update_small<0>, (const void*)k_update_small<1>,
                         (const void*)k_update_small<2>};
  if (g.variant) {   // option group: the option kernel, four calls only
    const void* kx = lds_cards ? (const void*)k_update_x<true> : (const void*)k_update_x<false>;
    if (fused) return fail(XA_ERR_STATE, "bsim4: option groups have no fused load");
    return launch_model_kernel(g, a, false, kx, kx);
  }
  if (g.persist_blocks > 0 && !fused) {
    const void* kp[2][3] = {
        {(const void*)k_update_persist<false, 0>, (const void*)k_update_persist<false, 1>,
         (const void*)k_update_persist<false, 2>},
        {(const void*)k_update_persist<true, 0>, (const void*)k_update_persist<true, 1>,
         (const void*)k_update_persist<true, 2>}};
    const int64_t tiles = grid_for(g.n, g.split_block);
    const unsigned grid = (unsigned)std::min<int64_t>(tiles, g.persist_blocks);
    void* args[] = {(void*)&a};
    XA_HIP(hipLaunchKernel(kp[lds_cards][cm], dim3(grid), dim3(g.split_block), args, 0,
                           g.ctx->stream));
    return XA_OK;
  }
  const void* k_stored = lds_cards && g.n <= kSmallUpdate ? small[cm] : stored[lds_cards][cm];
  return launch_model_kernel(g, a, fused, k_stored, fused_k[fused_lds][cm]);
}

// Workgroups of the persistent update resident on the whole GPU at its
// occupancy (the default groups' form; option groups keep one per tile).
static int persistent_capacity(const Group& g) {
  if (g.variant) return 0;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const bool lds = g.card_tables && g.card_block == g.split_block;
  const void* k = lds ? (const void*)k_update_persist<true, 0> : (const void*)k_update_persist<false, 0>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, g.split_block, 0) != hipSuccess)
    return 0;
  return cus * per_cu;
}

static int launch_pattern(Group& g, const xa_solver_state& ss, const xa_vectors& v) {
  if (!v.nextSol || !v.nextSto) return fail(XA_ERR_ARG, "bsim4 update: nextSol/nextSto required");
  if (ss.newtonIter == 0 && (!ss.dcopFlag || (ss.locaEnabledFlag && ss.dcopFlag)) && !v.currSto)
    return fail(XA_ERR_ARG, "bsim4 update: currSto required");
  if (ss.inputOPFlag && ss.initJctFlag && !v.flagSol)
    return fail(XA_ERR_ARG, "bsim4 update: flagSol required with inputOPFlag");
  KArgs a;
  fill_kargs(g, ss, v, true, a);   // kinds with an output, a target and not skipped
  const bool lds_cards = g.card_tables && g.card_block == g.pattern_B &&
                         a.mstride == kCardStride;
  const void* k[2][3] = {
      {(const void*)k_update_pattern<false, 0>, (const void*)k_update_pattern<false, 1>,
       (const void*)k_update_pattern<false, 2>},
      {(const void*)k_update_pattern<true, 0>, (const void*)k_update_pattern<true, 1>,
       (const void*)k_update_pattern<true, 2>}};
  return launch_pattern_kernel(g, a, k[lds_cards][card_mode(g)]);
}

// Algorithmic bytes per instance: the SoA columns and card index read once,
// node voltages (6 LIDs + values), 14 store LIDs, 12 store writes, 3 state
// LIDs + writes, vgb read + write, junction charge state written; the loads
// write value + target per entry.
static double traffic(const Group& g, int phase) {
  // per-instance columns read: those outside the card mode's set (the others
  // come with the card, one copy per model)
  const double cols = ND - card_columns(g);
  const double lids = 6 + NSTO + NSTA_BASE + 1 /*model index*/;
  if (phase == 0)   // transient: the junction charge state is written, not read
    return 8.0 * (cols + 6 /*node values*/ + 12 /*sto*/ + 3 /*sta*/ + 2 /*vgb*/ + 4 /*qb,cap*/) + 4.0 * lids;
  const bool lim = g.opt.voltageLimiterFlag != 0;
  if (phase == 1) return 16.0 * (4 /*F*/ + 4 /*Q*/ + (lim ? 8 : 0));
  return 16.0 * (38 + 16);
}

}  // namespace bsim4

const ModelImpl* model_bsim4() {
  static ModelImpl m = [] {
    using namespace bsim4;
    for (auto& d : g_mparams) d.is_temp = std::strcmp(d.name, "TNOM") == 0;
    for (auto& d : g_iparams) d.is_temp = std::strcmp(d.name, "TEMP") == 0;
    ModelImpl x{};
    x.type = XA_BSIM4;
    x.name = "bsim4";
    x.mparams = g_mparams;
    x.n_mparams = sizeof(g_mparams) / sizeof(ParamDef);
    x.n_mslots = NM;
    x.iparams = g_iparams;
    x.n_iparams = sizeof(g_iparams) / sizeof(ParamDef);
    x.n_islots = NI;
    x.max_nodes = NNODE;
    x.max_sto = NSTO;
    x.max_sta = NSTA;
    const int E[NK] = {8, 4, 4, 4, 38, 16, 0, 0};
    for (int k = 0; k < NK; ++k) x.max_entries[k] = E[k];
    x.setup = setup;
    x.resolve_nodes = resolve_nodes;
    x.stamps = stamps;
    x.launch_update = launch_update;
    x.launch_secondary = nullptr;
    x.skip_kind = skip_kind;
    x.state_names = kStateNames;
    x.n_state = NS;
    x.card_columns = card_columns;
    x.persistent_capacity = persistent_capacity;
    x.traffic = traffic;
    x.fused_block = 128;
    x.launch_pattern = launch_pattern;
    x.fused_at_scale = false;
    // the four calls beat the fused kernel at every size: one block's
    // in-order sums over shared targets are serial in the fused writeout
    x.fused_small = false;
    return x;
  }();
  return &m;
}

}  // namespace xa
