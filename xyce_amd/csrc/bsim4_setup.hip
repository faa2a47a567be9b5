// bsim4_setup.hip -- BSIM4 (level 54) host setup: model card defaults and
// processParams, instance defaults and processParams, updateTemperature
// with the size-dependent (binned) parameters.  Host code only; the
// evaluation kernel lives in model_bsim4.hip.
//
// Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET_B4.C
//   Model ctor defaults :15661-16056, Model::processParams :14740-14780,
//   Instance ctor defaults :4585-4618, Instance::processParams :3621-3794,
//   updateTemperature :5640-7734, DioIjthVjmEval :14292-14303,
//   NumFingerDiff :14322-14351, PAeffGeo :14364-14462,
//   RdseffGeo :14473-14588, RdsEndIso :14599-14657, RdsEndSha :14668-14726.
// Every option of the evaluation is taken; capMod, igcMod and igbMod values
// the evaluation has no branch for are rejected with XA_ERR_UNSUPPORTED.
#include <cmath>
#include <cstring>
#include "bsim4_host.h"

// The host setup (processParams, updateTemperature) keeps the C library's
// exp/log/pow in every build, as the CPU restatement's BSIM4 setup does.
#undef exp
#undef log
#undef pow
#undef tan
#undef sin

namespace xa {
namespace bsim4 {

static void DioIjthVjmEval(double Nvtm, double Ijth, double Isb, double XExpBV, double* Vjm) {
  double Tb, Tc, EVjmovNv;
  Tc = XExpBV;
  Tb = 1.0 + Ijth / Isb - Tc;
  EVjmovNv = 0.5 * (Tb + sqrt(Tb * Tb + 4.0 * Tc));
  *Vjm = Nvtm * log(EVjmovNv);
}

static void NumFingerDiff(double nf_arg, int minSD, double* nuIntD, double* nuEndD,
                          double* nuIntS, double* nuEndS) {
  int NF = (int)nf_arg;
  if ((NF % 2) != 0) {
    *nuEndD = *nuEndS = 1.0;
    *nuIntD = *nuIntS = 2.0 * dmax((nf_arg - 1.0) / 2.0, 0.0);
  } else {
    if (minSD == 1) {
      *nuEndD = 2.0;
      *nuIntD = 2.0 * dmax((nf_arg / 2.0 - 1.0), 0.0);
      *nuEndS = 0.0;
      *nuIntS = nf_arg;
    } else {
      *nuEndD = 0.0;
      *nuIntD = nf_arg;
      *nuEndS = 2.0;
      *nuIntS = 2.0 * dmax((nf_arg / 2.0 - 1.0), 0.0);
    }
  }
}

static void PAeffGeo(double nf_arg, int geo, int minSD, double Weffcj, double DMCG, double DMCI,
                     double DMDG, double* Ps, double* Pd, double* As, double* Ad) {
  double T0, T1, T2, ADiso, ADsha, ADmer, ASiso, ASsha, ASmer;
  double PDiso, PDsha, PDmer, PSiso, PSsha, PSmer;
  double nuIntD = 0.0, nuEndD = 0.0, nuIntS = 0.0, nuEndS = 0.0;
  if (geo < 9) NumFingerDiff(nf_arg, minSD, &nuIntD, &nuEndD, &nuIntS, &nuEndS);
  T0 = DMCG + DMCI;
  T1 = DMCG + DMCG;
  T2 = DMDG + DMDG;
  PSiso = PDiso = T0 + T0 + Weffcj;
  PSsha = PDsha = T1;
  PSmer = PDmer = T2;
  ASiso = ADiso = T0 * Weffcj;
  ASsha = ADsha = DMCG * Weffcj;
  ASmer = ADmer = DMDG * Weffcj;
  switch (geo) {
    case 0: *Ps = nuEndS * PSiso + nuIntS * PSsha; *Pd = nuEndD * PDiso + nuIntD * PDsha;
            *As = nuEndS * ASiso + nuIntS * ASsha; *Ad = nuEndD * ADiso + nuIntD * ADsha; break;
    case 1: *Ps = nuEndS * PSiso + nuIntS * PSsha; *Pd = (nuEndD + nuIntD) * PDsha;
            *As = nuEndS * ASiso + nuIntS * ASsha; *Ad = (nuEndD + nuIntD) * ADsha; break;
    case 2: *Ps = (nuEndS + nuIntS) * PSsha; *Pd = nuEndD * PDiso + nuIntD * PDsha;
            *As = (nuEndS + nuIntS) * ASsha; *Ad = nuEndD * ADiso + nuIntD * ADsha; break;
    case 3: *Ps = (nuEndS + nuIntS) * PSsha; *Pd = (nuEndD + nuIntD) * PDsha;
            *As = (nuEndS + nuIntS) * ASsha; *Ad = (nuEndD + nuIntD) * ADsha; break;
    case 4: *Ps = nuEndS * PSiso + nuIntS * PSsha; *Pd = nuEndD * PDmer + nuIntD * PDsha;
            *As = nuEndS * ASiso + nuIntS * ASsha; *Ad = nuEndD * ADmer + nuIntD * ADsha; break;
    case 5: *Ps = (nuEndS + nuIntS) * PSsha; *Pd = nuEndD * PDmer + nuIntD * PDsha;
            *As = (nuEndS + nuIntS) * ASsha; *Ad = nuEndD * ADmer + nuIntD * ADsha; break;
    case 6: *Ps = nuEndS * PSmer + nuIntS * PSsha; *Pd = nuEndD * PDiso + nuIntD * PDsha;
            *As = nuEndS * ASmer + nuIntS * ASsha; *Ad = nuEndD * ADiso + nuIntD * ADsha; break;
    case 7: *Ps = nuEndS * PSmer + nuIntS * PSsha; *Pd = (nuEndD + nuIntD) * PDsha;
            *As = nuEndS * ASmer + nuIntS * ASsha; *Ad = (nuEndD + nuIntD) * ADsha; break;
    case 8: *Ps = nuEndS * PSmer + nuIntS * PSsha; *Pd = nuEndD * PDmer + nuIntD * PDsha;
            *As = nuEndS * ASmer + nuIntS * ASsha; *Ad = nuEndD * ADmer + nuIntD * ADsha; break;
    case 9: *Ps = PSiso + (nf_arg - 1.0) * PSsha; *Pd = nf_arg * PDsha;
            *As = ASiso + (nf_arg - 1.0) * ASsha; *Ad = nf_arg * ADsha; break;
    case 10: *Ps = nf_arg * PSsha; *Pd = PDiso + (nf_arg - 1.0) * PDsha;
             *As = nf_arg * ASsha; *Ad = ADiso + (nf_arg - 1.0) * ADsha; break;
    default: break;
  }
}

static void RdsEndIso(double Weffcj, double Rsh, double DMCG, double DMCI, double DMDG,
                      double nuEnd, int rgeo, int Type, double* Rend) {
  (void)DMDG;
  if (Type == 1) {
    switch (rgeo) {
      case 1: case 2: case 5:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
      case 3: case 4: case 6:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * Weffcj / (3.0 * nuEnd * (DMCG + DMCI)); break;
      default: break;
    }
  } else {
    switch (rgeo) {
      case 1: case 3: case 7:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
      case 2: case 4: case 8:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * Weffcj / (3.0 * nuEnd * (DMCG + DMCI)); break;
      default: break;
    }
  }
}

/* The reference declares RdsEndSha(..., int rgeo, int Type, double nuEnd, ...)
 * but its callers pass (..., (int)nuEnd, rgeo, Type, ...) (N_DEV_MOSFET_B4.C
 * :14512 vs :14668); the parameters are restated in declaration order and
 * called the same way, so the argument shift is preserved. */
static void RdsEndSha(double Weffcj, double Rsh, double DMCG, double DMCI, double DMDG,
                      int rgeo, int Type, double nuEnd, double* Rend) {
  (void)DMCI; (void)DMDG;
  if (Type == 1) {
    switch (rgeo) {
      case 1: case 2: case 5:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
      case 3: case 4: case 6:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * Weffcj / (6.0 * nuEnd * DMCG); break;
      default: break;
    }
  } else {
    switch (rgeo) {
      case 1: case 3: case 7:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
      case 2: case 4: case 8:
        *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * Weffcj / (6.0 * nuEnd * DMCG); break;
      default: break;
    }
  }
}

static void RdseffGeo(double nf_arg, int geo, int rgeo, int minSD, double Weffcj, double Rsh,
                      double DMCG, double DMCI, double DMDG, int Type, double* Rtot) {
  double Rint = 0.0, Rend = 0.0;
  double nuIntD = 0.0, nuEndD = 0.0, nuIntS = 0.0, nuEndS = 0.0;
  if (geo < 9) {
    NumFingerDiff(nf_arg, minSD, &nuIntD, &nuEndD, &nuIntS, &nuEndS);
    if (Type == 1) Rint = (nuIntS == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuIntS);
    else Rint = (nuIntD == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuIntD);
  }
  switch (geo) {
    case 0:
      if (Type == 1) RdsEndIso(Weffcj, Rsh, DMCG, DMCI, DMDG, nuEndS, rgeo, 1, &Rend);
      else RdsEndIso(Weffcj, Rsh, DMCG, DMCI, DMDG, nuEndD, rgeo, 0, &Rend);
      break;
    case 1:
      if (Type == 1) RdsEndIso(Weffcj, Rsh, DMCG, DMCI, DMDG, nuEndS, rgeo, 1, &Rend);
      else RdsEndSha(Weffcj, Rsh, DMCG, DMCI, DMDG, (int)nuEndD, rgeo, 0, &Rend);
      break;
    case 2:
      if (Type == 1) RdsEndSha(Weffcj, Rsh, DMCG, DMCI, DMDG, (int)nuEndS, rgeo, 1, &Rend);
      else RdsEndIso(Weffcj, Rsh, DMCG, DMCI, DMDG, nuEndD, rgeo, 0, &Rend);
      break;
    case 3:
      if (Type == 1) RdsEndSha(Weffcj, Rsh, DMCG, DMCI, DMDG, (int)nuEndS, rgeo, 1, &Rend);
      else RdsEndSha(Weffcj, Rsh, DMCG, DMCI, DMDG, (int)nuEndD, rgeo, 0, &Rend);
      break;
    case 4:
      if (Type == 1) RdsEndIso(Weffcj, Rsh, DMCG, DMCI, DMDG, nuEndS, rgeo, 1, &Rend);
      else Rend = Rsh * DMDG / Weffcj;
      break;
    case 5:
      if (Type == 1) RdsEndSha(Weffcj, Rsh, DMCG, DMCI, DMDG, (int)nuEndS, rgeo, 1, &Rend);
      else Rend = Rsh * DMDG / (Weffcj * nuEndD);
      break;
    case 6:
      if (Type == 1) Rend = Rsh * DMDG / Weffcj;
      else RdsEndIso(Weffcj, Rsh, DMCG, DMCI, DMDG, nuEndD, rgeo, 0, &Rend);
      break;
    case 7:
      if (Type == 1) Rend = Rsh * DMDG / (Weffcj * nuEndS);
      else RdsEndSha(Weffcj, Rsh, DMCG, DMCI, DMDG, (int)nuEndD, rgeo, 0, &Rend);
      break;
    case 8:
      Rend = Rsh * DMDG / Weffcj;
      break;
    case 9:
      if (Type == 1) {
        Rend = 0.5 * Rsh * DMCG / Weffcj;
        Rint = (nf_arg == 2.0) ? 0.0 : Rsh * DMCG / (Weffcj * (nf_arg - 2.0));
      } else {
        Rend = 0.0;
        Rint = Rsh * DMCG / (Weffcj * nf_arg);
      }
      break;
    case 10:
      if (Type == 1) {
        Rend = 0.0;
        Rint = Rsh * DMCG / (Weffcj * nf_arg);
      } else {
        Rend = 0.5 * Rsh * DMCG / Weffcj;
        Rint = (nf_arg == 2.0) ? 0.0 : Rsh * DMCG / (Weffcj * (nf_arg - 2.0));
      }
      break;
    default: break;
  }
  if (Rint <= 0.0) *Rtot = Rend;
  else if (Rend <= 0.0) *Rtot = Rint;
  else *Rtot = Rint * Rend / (Rint + Rend);
}

/* ---------------------------------------------------------------------- */
/* Model constructor defaults and processParams (N_DEV_MOSFET_B4.C:15661-16056,
 * :14740-14780).  `c` holds the card in table order with given flags. */
int model_init(Model& mref, const Card& c, const xa_device_options& opt) {
  Model* m = &mref;
  int k = 0;
#define B4M(NAME, mem, def) m->p.mem = c.v[k]; m->g.mem = c.given[k]; ++k;
#include "xa_bsim4_params.def"
#undef B4M
  /* pseudo-parameter TYPE follows the table: the .model nmos/pmos keyword */
  m->dtype = (c.v[k] < 0) ? -1 : 1;
  if (!m->g.tnom) m->p.tnom = opt.tnom;
  if (!m->g.toxp) m->p.toxp = m->p.toxe;
  if (!m->g.toxm) m->p.toxm = m->p.toxe;
  if (!m->g.dsub) m->p.dsub = m->p.drout;
  if (!m->g.vth0) m->p.vth0 = (m->dtype == 1) ? 0.7 : -0.7;
  if (!m->g.vddeot) m->p.vddeot = (m->dtype == 1) ? 1.5 : -1.5;
  if (!m->g.eu) m->p.eu = (m->dtype == 1) ? 1.67 : 1.0;
  if (!m->g.ua) m->p.ua = (m->p.mobMod == 2) ? 1.0E-15 : 1.0E-9;
  if (!m->g.uc) m->p.uc = (m->p.mobMod == 1) ? -0.0465 : -0.0465E-9;
  if (!m->g.uc1) m->p.uc1 = (m->p.mobMod == 1) ? -0.056 : -0.056E-9;
  if (!m->g.u0) m->p.u0 = (m->dtype == 1) ? 0.067 : 0.025;
  if (!m->g.aigc) m->p.aigc = (m->dtype == 1) ? 1.36E-2 : 9.80E-3;
  if (!m->g.bigc) m->p.bigc = (m->dtype == 1) ? 1.71E-3 : 7.59E-4;
  if (!m->g.cigc) m->p.cigc = (m->dtype == 1) ? 0.075 : 0.03;
  if (m->g.aigsd) {
    m->p.aigs = m->p.aigd = m->p.aigsd;
  } else {
    m->p.aigsd = (m->dtype == 1) ? 1.36E-2 : 9.80E-3;
    if (!m->g.aigs) m->p.aigs = m->p.aigsd;
    if (!m->g.aigd) m->p.aigd = m->p.aigsd;
  }
  if (m->g.bigsd) {
    m->p.bigs = m->p.bigd = m->p.bigsd;
  } else {
    m->p.bigsd = (m->dtype == 1) ? 1.71E-3 : 7.59E-4;
    if (!m->g.bigs) m->p.bigs = m->p.bigsd;
    if (!m->g.bigd) m->p.bigd = m->p.bigsd;
  }
  if (m->g.cigsd) {
    m->p.cigs = m->p.cigd = m->p.cigsd;
  } else {
    m->p.cigsd = (m->dtype == 1) ? 0.075 : 0.03;
    if (!m->g.cigs) m->p.cigs = m->p.cigsd;
    if (!m->g.cigd) m->p.cigd = m->p.cigsd;
  }
  if (!m->g.ijthdfwd) m->p.ijthdfwd = m->p.ijthsfwd;
  if (!m->g.ijthdrev) m->p.ijthdrev = m->p.ijthsrev;
  if (!m->g.xjbvd) m->p.xjbvd = m->p.xjbvs;
  if (!m->g.bvd) m->p.bvd = m->p.bvs;
  if (!m->g.ckappad) m->p.ckappad = m->p.ckappas;
  if (!m->g.dmci) m->p.dmci = m->p.dmcg;
  if (!m->g.Llc) m->p.Llc = m->p.Ll;
  if (!m->g.Lwc) m->p.Lwc = m->p.Lw;
  if (!m->g.Lwlc) m->p.Lwlc = m->p.Lwl;
  if (!m->g.Wlc) m->p.Wlc = m->p.Wl;
  if (!m->g.Wwc) m->p.Wwc = m->p.Ww;
  if (!m->g.Wwlc) m->p.Wwlc = m->p.Wwl;
  if (!m->g.dwc) m->p.dwc = m->p.Wint;
  if (!m->g.dlc) m->p.dlc = m->p.Lint;
  if (!m->g.agisl) { if (m->g.agidl) m->p.agisl = m->p.agidl; }
  if (!m->g.bgisl) { if (m->g.bgidl) m->p.bgisl = m->p.bgidl; }
  if (!m->g.cgisl) { if (m->g.cgidl) m->p.cgisl = m->p.cgidl; }
  if (!m->g.egisl) { if (m->g.egidl) m->p.egisl = m->p.egidl; }
  if (!m->g.dlcig) m->p.dlcig = m->p.Lint;
  if (!m->g.dlcigd) {
    if (!m->g.dlcig) m->p.dlcigd = m->p.Lint;
    else m->p.dlcigd = m->p.dlcig;
  }
  if (!m->g.dwj) m->p.dwj = m->p.dwc;
  if (!m->g.cf) m->p.cf = 2.0 * m->p.epsrox * (kEps0 / 3.14159265358979323846) * log(1.0 + 0.4E-6 / m->p.toxe);
  if (!m->g.DjctSatCurDensity) m->p.DjctSatCurDensity = m->p.SjctSatCurDensity;
  if (!m->g.DjctSidewallSatCurDensity) m->p.DjctSidewallSatCurDensity = m->p.SjctSidewallSatCurDensity;
  if (!m->g.DjctGateSidewallSatCurDensity) m->p.DjctGateSidewallSatCurDensity = m->p.SjctGateSidewallSatCurDensity;
  if (!m->g.DbulkJctPotential) m->p.DbulkJctPotential = m->p.SbulkJctPotential;
  if (!m->g.DjctEmissionCoeff) m->p.DjctEmissionCoeff = m->p.SjctEmissionCoeff;
  if (!m->g.DjctTempExponent) m->p.DjctTempExponent = m->p.SjctTempExponent;
  if (!m->g.DbulkJctBotGradingCoeff) m->p.DbulkJctBotGradingCoeff = m->p.SbulkJctBotGradingCoeff;
  if (!m->g.DbulkJctSideGradingCoeff) m->p.DbulkJctSideGradingCoeff = m->p.SbulkJctSideGradingCoeff;
  if (!m->g.SbulkJctGateSideGradingCoeff) m->p.SbulkJctGateSideGradingCoeff = m->p.SbulkJctSideGradingCoeff;
  if (!m->g.DbulkJctGateSideGradingCoeff) m->p.DbulkJctGateSideGradingCoeff = m->p.SbulkJctGateSideGradingCoeff;
  if (!m->g.DsidewallJctPotential) m->p.DsidewallJctPotential = m->p.SsidewallJctPotential;
  if (!m->g.SGatesidewallJctPotential) m->p.SGatesidewallJctPotential = m->p.SsidewallJctPotential;
  if (!m->g.DGatesidewallJctPotential) m->p.DGatesidewallJctPotential = m->p.DsidewallJctPotential;
  if (!m->g.DunitAreaJctCap) m->p.DunitAreaJctCap = m->p.SunitAreaJctCap;
  if (!m->g.DunitLengthSidewallJctCap) m->p.DunitLengthSidewallJctCap = m->p.SunitLengthSidewallJctCap;
  if (!m->g.SunitLengthGateSidewallJctCap) m->p.SunitLengthGateSidewallJctCap = m->p.SunitLengthSidewallJctCap;
  if (!m->g.DunitLengthGateSidewallJctCap) m->p.DunitLengthGateSidewallJctCap = m->p.SunitLengthGateSidewallJctCap;
  if (!m->g.jtsd) m->p.jtsd = m->p.jtss;
  if (!m->g.jtsswd) m->p.jtsswd = m->p.jtssws;
  if (!m->g.jtsswgd) m->p.jtsswgd = m->p.jtsswgs;
  if (!m->g.njtsd) { if (m->g.njts) m->p.njtsd = m->p.njts; }
  if (!m->g.njtsswd) { if (m->g.njtssw) m->p.njtsswd = m->p.njtssw; }
  if (!m->g.njtsswgd) { if (m->g.njtsswg) m->p.njtsswgd = m->p.njtsswg; }
  if (!m->g.xtsd) m->p.xtsd = m->p.xtss;
  if (!m->g.xtsswd) m->p.xtsswd = m->p.xtssws;
  if (!m->g.xtsswgd) m->p.xtsswgd = m->p.xtsswgs;
  if (!m->g.tnjtsd) { if (m->g.tnjts) m->p.tnjtsd = m->p.tnjts; }
  if (!m->g.tnjtsswd) { if (m->g.tnjtssw) m->p.tnjtsswd = m->p.tnjtssw; }
  if (!m->g.tnjtsswgd) { if (m->g.tnjtsswg) m->p.tnjtsswgd = m->p.tnjtsswg; }
  if (!m->g.vtsd) m->p.vtsd = m->p.vtss;
  if (!m->g.vtsswd) m->p.vtsswd = m->p.vtssws;
  if (!m->g.vtsswgd) m->p.vtsswgd = m->p.vtsswgs;
  if (!m->g.lagisl) { if (m->g.lagidl) m->p.lagisl = m->p.lagidl; }
  if (!m->g.lbgisl) { if (m->g.lbgidl) m->p.lbgisl = m->p.lbgidl; }
  if (!m->g.lcgisl) { if (m->g.lcgidl) m->p.lcgisl = m->p.lcgidl; }
  if (!m->g.legisl) { if (m->g.legidl) m->p.legisl = m->p.legidl; }
  if (!(!m->g.aigsd && (m->g.aigs || m->g.aigd))) m->p.laigs = m->p.laigd = m->p.laigsd;
  if (!(!m->g.bigsd && (m->g.bigs || m->g.bigd))) m->p.lbigs = m->p.lbigd = m->p.lbigsd;
  if (!(!m->g.cigsd && (m->g.cigs || m->g.cigd))) m->p.lcigs = m->p.lcigd = m->p.lcigsd;
  if (!m->g.wagisl) { if (m->g.wagidl) m->p.wagisl = m->p.wagidl; }
  if (!m->g.wbgisl) { if (m->g.wbgidl) m->p.wbgisl = m->p.wbgidl; }
  if (!m->g.wcgisl) { if (m->g.wcgidl) m->p.wcgisl = m->p.wcgidl; }
  if (!m->g.wegisl) { if (m->g.wegidl) m->p.wegisl = m->p.wegidl; }
  if (!(!m->g.aigsd && (m->g.aigs || m->g.aigd))) m->p.waigs = m->p.waigd = m->p.waigsd;
  if (!(!m->g.bigsd && (m->g.bigs || m->g.bigd))) m->p.wbigs = m->p.wbigd = m->p.wbigsd;
  if (!(!m->g.cigsd && (m->g.cigs || m->g.cigd))) m->p.wcigs = m->p.wcigd = m->p.wcigsd;
  if (!m->g.pagisl) { if (m->g.pagidl) m->p.pagisl = m->p.pagidl; }
  if (!m->g.pbgisl) { if (m->g.pbgidl) m->p.pbgisl = m->p.pbgidl; }
  if (!m->g.pcgisl) { if (m->g.pcgidl) m->p.pcgisl = m->p.pcgidl; }
  if (!m->g.pegisl) { if (m->g.pegidl) m->p.pegisl = m->p.pegidl; }
  if (!(!m->g.aigsd && (m->g.aigs || m->g.aigd))) m->p.paigs = m->p.paigd = m->p.paigsd;
  if (!(!m->g.bigsd && (m->g.bigs || m->g.bigd))) m->p.pbigs = m->p.pbigd = m->p.pbigsd;
  if (!(!m->g.cigsd && (m->g.cigs || m->g.cigd))) m->p.pcigs = m->p.pcigd = m->p.pcigsd;
  /* Model::processParams */
  if (m->p.SbulkJctPotential < 0.1) m->p.SbulkJctPotential = 0.1;
  if (m->p.SsidewallJctPotential < 0.1) m->p.SsidewallJctPotential = 0.1;
  if (m->p.SGatesidewallJctPotential < 0.1) m->p.SGatesidewallJctPotential = 0.1;
  if (m->p.DbulkJctPotential < 0.1) m->p.DbulkJctPotential = 0.1;
  if (m->p.DsidewallJctPotential < 0.1) m->p.DsidewallJctPotential = 0.1;
  if (m->p.DGatesidewallJctPotential < 0.1) m->p.DGatesidewallJctPotential = 0.1;
  /* option subset implemented by this restatement */
  // updateTemperature rejects other junction models (N_DEV_MOSFET_B4.C:7515)
  if (m->p.dioMod != 0 && m->p.dioMod != 1 && m->p.dioMod != 2)
    return fail(XA_ERR_ARG, "bsim4: specified dioMod not matched");
  // the evaluation has branches for these values only (any other would
  // leave its charges or gate-current terms unset)
  if (m->p.capMod != 0 && m->p.capMod != 1 && m->p.capMod != 2)
    return fail(XA_ERR_UNSUPPORTED, "bsim4: capMod must be 0, 1 or 2");
  if (m->p.igcMod != 0 && m->p.igcMod != 1 && m->p.igcMod != 2)
    return fail(XA_ERR_UNSUPPORTED, "bsim4: igcMod must be 0, 1 or 2");
  if (m->p.igbMod != 0 && m->p.igbMod != 1) return fail(XA_ERR_UNSUPPORTED, "bsim4: igbMod must be 0 or 1");
  return XA_OK;
}


/* Instance::updateTemperature (N_DEV_MOSFET_B4.C:5640-7734) */
static void update_temperature(Inst* I, double temp_tmp, const xa_device_options& opt) {
  Model* m = I->model;
  Size* p;
  double tmp, tmp1, tmp2, Eg, Eg0, ni, epssub, T0, T1, T2, T3, T4, T5, T6, T8, T9, Lnew, Wnew;
  double delTemp, TRatio, Inv_L, Inv_W, Inv_LW, Vtm0, Tnom, dumPs, dumPd, dumAs, dumAd, PowWeffWr;
  double Nvtms, Nvtmd, SourceSatCurrent, DrainSatCurrent, T10;
  double Inv_saref, Inv_sbref, Inv_sa, Inv_sb, rho, Ldrn, dvth0_lod, W_tmp, Inv_ODeff, OD_offset;
  double dk2_lod, deta0_lod, lnl, lnw, lnnf, rbpbx, rbpby, rbsbx, rbsby, rbdbx, rbdby;
  double kvsat, wlod, sceff, Wdrn, toxe, epsrox, T7;
  int bodymode;
  const double nf = I->nf;
  (void)opt;
  if (temp_tmp != -999.0) I->temp = temp_tmp;
  Tnom = m->p.tnom;
  TRatio = I->temp / Tnom;
  if (m->p.mtrlMod == 0) {
    if ((m->g.toxe) && (m->g.toxp) && (m->g.dtox) && (m->p.toxe != (m->p.toxp + m->p.dtox))) {
    } else if ((m->g.toxe) && (!m->g.toxp)) {
      m->p.toxp = m->p.toxe - m->p.dtox;
    } else if ((!m->g.toxe) && (m->g.toxp)) {
      m->p.toxe = m->p.toxp + m->p.dtox;
    }
  }
  if (m->p.mtrlMod) {
    epsrox = 3.9;
    toxe = m->p.eot;
    epssub = kEps0 * m->p.epsrsub;
  } else {
    epsrox = m->p.epsrox;
    toxe = m->p.toxe;
    epssub = kEpsSi;
  }
  m->coxe = epsrox * kEps0 / toxe;
  if (m->p.mtrlMod == 0) m->coxp = m->p.epsrox * kEps0 / m->p.toxp;
  if (!m->g.cgdo) {
    if (m->g.dlc && (m->p.dlc > 0.0)) m->p.cgdo = m->p.dlc * m->coxe - m->p.cgdl;
    else m->p.cgdo = 0.6 * m->p.xj * m->coxe;
  }
  if (!m->g.cgso) {
    if (m->g.dlc && (m->p.dlc > 0.0)) m->p.cgso = m->p.dlc * m->coxe - m->p.cgsl;
    else m->p.cgso = 0.6 * m->p.xj * m->coxe;
  }
  if (!m->g.cgbo) m->p.cgbo = 2.0 * m->p.dwc * m->coxe;
  m->vcrit = kVt0 * log(kVt0 / (sqrt(2.0) * 1.0e-14));
  m->factor1 = sqrt(epssub / (epsrox * kEps0) * toxe);
  Vtm0 = m->vtm0 = kKboQ * Tnom;
  if (m->p.mtrlMod == 0) {
    Eg0 = 1.16 - 7.02e-4 * Tnom * Tnom / (Tnom + 1108.0);
    ni = 1.45e10 * (Tnom / 300.15) * sqrt(Tnom / 300.15) * exp(21.5565981 - Eg0 / (2.0 * Vtm0));
  } else {
    Eg0 = m->p.bg0sub - m->p.tbgasub * Tnom * Tnom / (Tnom + m->p.tbgbsub);
    T0 = m->p.bg0sub - m->p.tbgasub * 90090.0225 / (300.15 + m->p.tbgbsub);
    ni = m->p.ni0sub * (Tnom / 300.15) * sqrt(Tnom / 300.15) * exp((T0 - Eg0) / (2.0 * Vtm0));
  }
  m->Eg0 = Eg0;
  m->vtm = kKboQ * I->temp;
  if (m->p.mtrlMod == 0) Eg = 1.16 - 7.02e-4 * I->temp * I->temp / (I->temp + 1108.0);
  else Eg = m->p.bg0sub - m->p.tbgasub * I->temp * I->temp / (I->temp + m->p.tbgbsub);
  if (I->temp != Tnom) {
    T0 = Eg0 / Vtm0 - Eg / m->vtm;
    T1 = log(I->temp / Tnom);
    T2 = T0 + m->p.SjctTempExponent * T1;
    T3 = exp(T2 / m->p.SjctEmissionCoeff);
    m->SjctTempSatCurDensity = m->p.SjctSatCurDensity * T3;
    m->SjctSidewallTempSatCurDensity = m->p.SjctSidewallSatCurDensity * T3;
    m->SjctGateSidewallTempSatCurDensity = m->p.SjctGateSidewallSatCurDensity * T3;
    T2 = T0 + m->p.DjctTempExponent * T1;
    T3 = exp(T2 / m->p.DjctEmissionCoeff);
    m->DjctTempSatCurDensity = m->p.DjctSatCurDensity * T3;
    m->DjctSidewallTempSatCurDensity = m->p.DjctSidewallSatCurDensity * T3;
    m->DjctGateSidewallTempSatCurDensity = m->p.DjctGateSidewallSatCurDensity * T3;
  } else {
    m->SjctTempSatCurDensity = m->p.SjctSatCurDensity;
    m->SjctSidewallTempSatCurDensity = m->p.SjctSidewallSatCurDensity;
    m->SjctGateSidewallTempSatCurDensity = m->p.SjctGateSidewallSatCurDensity;
    m->DjctTempSatCurDensity = m->p.DjctSatCurDensity;
    m->DjctSidewallTempSatCurDensity = m->p.DjctSidewallSatCurDensity;
    m->DjctGateSidewallTempSatCurDensity = m->p.DjctGateSidewallSatCurDensity;
  }
  if (m->SjctTempSatCurDensity < 0.0) m->SjctTempSatCurDensity = 0.0;
  if (m->SjctSidewallTempSatCurDensity < 0.0) m->SjctSidewallTempSatCurDensity = 0.0;
  if (m->SjctGateSidewallTempSatCurDensity < 0.0) m->SjctGateSidewallTempSatCurDensity = 0.0;
  if (m->DjctTempSatCurDensity < 0.0) m->DjctTempSatCurDensity = 0.0;
  if (m->DjctSidewallTempSatCurDensity < 0.0) m->DjctSidewallTempSatCurDensity = 0.0;
  if (m->DjctGateSidewallTempSatCurDensity < 0.0) m->DjctGateSidewallTempSatCurDensity = 0.0;
  delTemp = I->temp - Tnom;
  T0 = m->p.tcj * delTemp;
  if (T0 >= -1.0) {
    m->SunitAreaTempJctCap = m->p.SunitAreaJctCap * (1.0 + T0);
    m->DunitAreaTempJctCap = m->p.DunitAreaJctCap * (1.0 + T0);
  } else {
    if (m->p.SunitAreaJctCap > 0.0) m->SunitAreaTempJctCap = 0.0;
    if (m->p.DunitAreaJctCap > 0.0) m->DunitAreaTempJctCap = 0.0;
  }
  T0 = m->p.tcjsw * delTemp;
  if (T0 >= -1.0) {
    m->SunitLengthSidewallTempJctCap = m->p.SunitLengthSidewallJctCap * (1.0 + T0);
    m->DunitLengthSidewallTempJctCap = m->p.DunitLengthSidewallJctCap * (1.0 + T0);
  } else {
    if (m->p.SunitLengthSidewallJctCap > 0.0) m->SunitLengthSidewallTempJctCap = 0.0;
    if (m->p.DunitLengthSidewallJctCap > 0.0) m->DunitLengthSidewallTempJctCap = 0.0;
  }
  T0 = m->p.tcjswg * delTemp;
  if (T0 >= -1.0) {
    m->SunitLengthGateSidewallTempJctCap = m->p.SunitLengthGateSidewallJctCap * (1.0 + T0);
    m->DunitLengthGateSidewallTempJctCap = m->p.DunitLengthGateSidewallJctCap * (1.0 + T0);
  } else {
    if (m->p.SunitLengthGateSidewallJctCap > 0.0) m->SunitLengthGateSidewallTempJctCap = 0.0;
    if (m->p.DunitLengthGateSidewallJctCap > 0.0) m->DunitLengthGateSidewallTempJctCap = 0.0;
  }
  m->PhiBS = m->p.SbulkJctPotential - m->p.tpb * delTemp;
  if (m->PhiBS < 0.01) m->PhiBS = 0.01;
  m->PhiBD = m->p.DbulkJctPotential - m->p.tpb * delTemp;
  if (m->PhiBD < 0.01) m->PhiBD = 0.01;
  m->PhiBSWS = m->p.SsidewallJctPotential - m->p.tpbsw * delTemp;
  if (m->PhiBSWS <= 0.01) m->PhiBSWS = 0.01;
  m->PhiBSWD = m->p.DsidewallJctPotential - m->p.tpbsw * delTemp;
  if (m->PhiBSWD <= 0.01) m->PhiBSWD = 0.01;
  m->PhiBSWGS = m->p.SGatesidewallJctPotential - m->p.tpbswg * delTemp;
  if (m->PhiBSWGS <= 0.01) m->PhiBSWGS = 0.01;
  m->PhiBSWGD = m->p.DGatesidewallJctPotential - m->p.tpbswg * delTemp;
  if (m->PhiBSWGD <= 0.01) m->PhiBSWGD = 0.01;
  if (m->p.ijthdfwd <= 0.0) m->p.ijthdfwd = 0.1;
  if (m->p.ijthsfwd <= 0.0) m->p.ijthsfwd = 0.1;
  if (m->p.ijthdrev <= 0.0) m->p.ijthdrev = 0.1;
  if (m->p.ijthsrev <= 0.0) m->p.ijthsrev = 0.1;
  if ((m->p.xjbvd <= 0.0) && (m->p.dioMod == 2)) m->p.xjbvd = 1.0;
  else if ((m->p.xjbvd < 0.0) && (m->p.dioMod == 0)) m->p.xjbvd = 1.0;
  if (m->p.bvd <= 0.0) m->p.bvd = 10.0;
  if ((m->p.xjbvs <= 0.0) && (m->p.dioMod == 2)) m->p.xjbvs = 1.0;
  else if ((m->p.xjbvs < 0.0) && (m->p.dioMod == 0)) m->p.xjbvs = 1.0;
  if (m->p.bvs <= 0.0) m->p.bvs = 10.0;

  Ldrn = I->l;
  Wdrn = I->w / nf;
  {
    auto it = m->sizes.find(std::make_tuple(I->l, I->w, nf));
    p = (it == m->sizes.end()) ? nullptr : it->second.get();
  }
  Lnew = I->l + m->p.xl;
  Wnew = I->w / nf + m->p.xw;
  if (p == nullptr) {
    std::unique_ptr<Size> fresh(new Size());
    p = fresh.get();
    m->sizes[std::make_tuple(I->l, I->w, nf)] = std::move(fresh);
    p->referenceTemperature = temp_tmp;
    p->Length = I->l;
    p->Width = I->w;
    p->NFinger = nf;
    T0 = pow(Lnew, m->p.Lln);
    T1 = pow(Wnew, m->p.Lwn);
    tmp1 = m->p.Ll / T0 + m->p.Lw / T1 + m->p.Lwl / (T0 * T1);
    p->dl = m->p.Lint + tmp1;
    tmp2 = m->p.Llc / T0 + m->p.Lwc / T1 + m->p.Lwlc / (T0 * T1);
    p->dlc = m->p.dlc + tmp2;
    T2 = pow(Lnew, m->p.Wln);
    T3 = pow(Wnew, m->p.Wwn);
    tmp1 = m->p.Wl / T2 + m->p.Ww / T3 + m->p.Wwl / (T2 * T3);
    p->dw = m->p.Wint + tmp1;
    tmp2 = m->p.Wlc / T2 + m->p.Wwc / T3 + m->p.Wwlc / (T2 * T3);
    p->dwc = m->p.dwc + tmp2;
    p->dwj = m->p.dwj + tmp2;
    p->leff = Lnew - 2.0 * p->dl;
    p->weff = Wnew - 2.0 * p->dw;
    p->leffCV = Lnew - 2.0 * p->dlc;
    p->weffCV = Wnew - 2.0 * p->dwc;
    p->weffCJ = Wnew - 2.0 * p->dwj;
    if (m->p.binUnit == 1) {
      Inv_L = 1.0e-6 / p->leff;
      Inv_W = 1.0e-6 / p->weff;
      Inv_LW = 1.0e-12 / (p->leff * p->weff);
    } else {
      Inv_L = 1.0 / p->leff;
      Inv_W = 1.0 / p->weff;
      Inv_LW = 1.0 / (p->leff * p->weff);
    }
#define B4BIN(sm, base, l, w, pp) p->sm = m->p.base + m->p.l * Inv_L + m->p.w * Inv_W + m->p.pp * Inv_LW;
#include "xa_bsim4_binning.def"
#undef B4BIN
    p->abulkCVfactor = 1.0 + pow((p->clc / p->leffCV), p->cle);
    T0 = (TRatio - 1.0);
    PowWeffWr = pow(p->weffCJ * 1.0e6, p->wr) * nf;
    T1 = T2 = T3 = T4 = 0.0;
    if (m->p.tempMod == 0) {
      p->ua = p->ua + p->ua1 * T0;
      p->ub = p->ub + p->ub1 * T0;
      p->uc = p->uc + p->uc1 * T0;
      p->ud = p->ud + p->ud1 * T0;
      p->vsattemp = p->vsat - p->at * T0;
      T10 = p->prt * T0;
      if (m->p.rdsMod) {
        T1 = p->rdw + T10;
        T2 = m->p.rdwmin + T10;
        T3 = p->rsw + T10;
        T4 = m->p.rswmin + T10;
      }
      p->rds0 = (p->rdsw + T10) * nf / PowWeffWr;
      p->rdswmin = (m->p.rdswmin + T10) * nf / PowWeffWr;
    } else {
      p->ua = p->ua * (1.0 + p->ua1 * delTemp);
      p->ub = p->ub * (1.0 + p->ub1 * delTemp);
      p->uc = p->uc * (1.0 + p->uc1 * delTemp);
      p->ud = p->ud * (1.0 + p->ud1 * delTemp);
      p->vsattemp = p->vsat * (1.0 - p->at * delTemp);
      T10 = 1.0 + p->prt * delTemp;
      if (m->p.rdsMod) {
        T1 = p->rdw * T10;
        T2 = m->p.rdwmin * T10;
        T3 = p->rsw * T10;
        T4 = m->p.rswmin * T10;
      }
      p->rds0 = p->rdsw * T10 * nf / PowWeffWr;
      p->rdswmin = m->p.rdswmin * T10 * nf / PowWeffWr;
    }
    if (T1 < 0.0) T1 = 0.0;
    if (T2 < 0.0) T2 = 0.0;
    p->rd0 = T1 / PowWeffWr;
    p->rdwmin = T2 / PowWeffWr;
    if (T3 < 0.0) T3 = 0.0;
    if (T4 < 0.0) T4 = 0.0;
    p->rs0 = T3 / PowWeffWr;
    p->rswmin = T4 / PowWeffWr;
    if (p->u0 > 1.0) p->u0 = p->u0 / 1.0e4;
    T5 = 1.0 - p->up * exp(-p->leff / p->lp);
    p->u0temp = p->u0 * T5 * pow(TRatio, p->ute);
    if (p->eu < 0.0) p->eu = 0.0;
    p->vfbsdoff = p->vfbsdoff * (1.0 + p->tvfbsdoff * delTemp);
    p->voff = p->voff * (1.0 + p->tvoff * delTemp);
    if ((m->g.vtl) && (m->p.vtl > 0.0)) {
      if (m->p.lc < 0.0) p->lc = 0.0;
      else p->lc = m->p.lc;
      T0 = p->leff / (p->xn * p->leff + p->lc);
      p->tfactor = (1.0 - T0) / (1.0 + T0);
    }
    p->cgdo = (m->p.cgdo + p->cf) * p->weffCV;
    p->cgso = (m->p.cgso + p->cf) * p->weffCV;
    p->cgbo = m->p.cgbo * p->leffCV * nf;
    if (!m->g.ndep && m->g.gamma1) {
      T0 = p->gamma1 * m->coxe;
      p->ndep = 3.01248e22 * T0 * T0;
    }
    p->phi = Vtm0 * log(p->ndep / ni) + p->phin + 0.4;
    p->sqrtPhi = sqrt(p->phi);
    p->phis3 = p->sqrtPhi * p->phi;
    p->Xdep0 = sqrt(2.0 * epssub / (kChargeQ * p->ndep * 1.0e6)) * p->sqrtPhi;
    p->sqrtXdep0 = sqrt(p->Xdep0);
    if (m->p.mtrlMod == 0) p->litl = sqrt(3.0 * p->xj * toxe);
    else p->litl = sqrt(m->p.epsrsub / epsrox * p->xj * toxe);
    p->vbi = Vtm0 * log(p->nsd * p->ndep / (ni * ni));
    if (m->p.mtrlMod == 0) {
      if (p->ngate > 0.0) p->vfbsd = Vtm0 * log(p->ngate / p->nsd);
      else p->vfbsd = 0.0;
    } else {
      T0 = Vtm0 * log(p->nsd / ni);
      T1 = 0.5 * Eg0;
      if (T0 > T1) T0 = T1;
      T2 = m->p.easub + T1 - m->dtype * T0;
      p->vfbsd = m->p.phig - T2;
    }
    p->cdep0 = sqrt(kChargeQ * epssub * p->ndep * 1.0e6 / 2.0 / p->phi);
    p->ToxRatio = exp(p->ntox * log(m->p.toxref / toxe)) / toxe / toxe;
    p->ToxRatioEdge = exp(p->ntox * log(m->p.toxref / (toxe * p->poxedge))) / toxe / toxe
                      / p->poxedge / p->poxedge;
    p->Aechvb = (m->dtype == 1) ? 4.97232e-7 : 3.42537e-7;
    p->Bechvb = (m->dtype == 1) ? 7.45669e11 : 1.16645e12;
    p->AechvbEdgeS = p->Aechvb * p->weff * m->p.dlcig * p->ToxRatioEdge;
    p->AechvbEdgeD = p->Aechvb * p->weff * m->p.dlcigd * p->ToxRatioEdge;
    p->BechvbEdge = -p->Bechvb * m->p.toxe * p->poxedge;
    p->Aechvb *= p->weff * p->leff * p->ToxRatio;
    p->Bechvb *= -toxe;
    p->mstar = 0.5 + atan(p->minv) / 3.14159265358979323846;
    p->mstarcv = 0.5 + atan(p->minvcv) / 3.14159265358979323846;
    p->voffcbn = p->voff + m->p.voffl / p->leff;
    p->voffcbncv = p->voffcv + m->p.voffcvl / p->leff;
    p->ldeb = sqrt(epssub * Vtm0 / (kChargeQ * p->ndep * 1.0e6)) / 3.0;
    p->acde *= pow((p->ndep / 2.0e16), -0.25);
    if (m->g.k1 || m->g.k2) {
      if (!m->g.k1) p->k1 = 0.53;
      if (!m->g.k2) p->k2 = -0.0186;
    } else {
      if (!m->g.vbx) p->vbx = p->phi - 7.7348e-4 * p->ndep * p->xt * p->xt;
      if (p->vbx > 0.0) p->vbx = -p->vbx;
      if (p->vbm > 0.0) p->vbm = -p->vbm;
      if (!m->g.gamma1) p->gamma1 = 5.753e-12 * sqrt(p->ndep) / m->coxe;
      if (!m->g.gamma2) p->gamma2 = 5.753e-12 * sqrt(p->nsub) / m->coxe;
      T0 = p->gamma1 - p->gamma2;
      T1 = sqrt(p->phi - p->vbx) - p->sqrtPhi;
      T2 = sqrt(p->phi * (p->phi - p->vbm)) - p->phi;
      p->k2 = T0 * T1 / (2.0 * T2 + p->vbm);
      p->k1 = p->gamma2 - 2.0 * p->k2 * sqrt(p->phi - p->vbm);
    }
    if (!m->g.vfb) {
      if (m->g.vth0) p->vfb = m->dtype * p->vth0 - p->phi - p->k1 * p->sqrtPhi;
      else if ((m->p.mtrlMod) && (m->g.phig) && (m->g.nsub)) {
        T0 = Vtm0 * log(p->nsub / ni);
        T1 = 0.5 * Eg0;
        if (T0 > T1) T0 = T1;
        T2 = m->p.easub + T1 + m->dtype * T0;
        p->vfb = m->p.phig - T2;
      } else {
        p->vfb = -1.0;
      }
    }
    if (!m->g.vth0) p->vth0 = m->dtype * (p->vfb + p->phi + p->k1 * p->sqrtPhi);
    p->k1ox = p->k1 * toxe / m->p.toxm;
    tmp = sqrt(epssub / (epsrox * kEps0) * toxe * p->Xdep0);
    T0 = p->dsub * p->leff / tmp;
    if (T0 < kExpThreshold) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * kMinExp;
      p->theta0vb0 = T1 / T4;
    } else {
      p->theta0vb0 = 1.0 / (kMaxExp - 2.0);
    }
    T0 = p->drout * p->leff / tmp;
    if (T0 < kExpThreshold) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * kMinExp;
      T5 = T1 / T4;
    } else {
      T5 = 1.0 / (kMaxExp - 2.0);
    }
    p->thetaRout = p->pdibl1 * T5 + p->pdibl2;
    tmp = sqrt(p->Xdep0);
    tmp1 = p->vbi - p->phi;
    tmp2 = m->factor1 * tmp;
    T0 = p->dvt1w * p->weff * p->leff / tmp2;
    if (T0 < kExpThreshold) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * kMinExp;
      T8 = T1 / T4;
    } else {
      T8 = 1.0 / (kMaxExp - 2.0);
    }
    T0 = p->dvt0w * T8;
    T8 = T0 * tmp1;
    T0 = p->dvt1 * p->leff / tmp2;
    if (T0 < kExpThreshold) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * kMinExp;
      T9 = T1 / T4;
    } else {
      T9 = 1.0 / (kMaxExp - 2.0);
    }
    T9 = p->dvt0 * T9 * tmp1;
    T4 = toxe * p->phi / (p->weff + p->w0);
    T0 = sqrt(1.0 + p->lpe0 / p->leff);
    T3 = 0.0;
    if ((m->p.tempMod == 1) || (m->p.tempMod == 0)) T3 = (p->kt1 + p->kt1l / p->leff) * (TRatio - 1.0);
    if (m->p.tempMod == 2) T3 = -p->kt1 * (TRatio - 1.0);
    T5 = p->k1ox * (T0 - 1.0) * p->sqrtPhi + T3;
    p->vfbzbfactor = -T8 - T9 + p->k3 * T4 + T5 - p->phi - p->k1 * p->sqrtPhi;
    wlod = m->p.wlod;
    if (m->p.wlod < 0.0) wlod = 0.0;
    T0 = pow(Lnew, m->p.llodku0);
    W_tmp = Wnew + wlod;
    T1 = pow(W_tmp, m->p.wlodku0);
    tmp1 = m->p.lku0 / T0 + m->p.wku0 / T1 + m->p.pku0 / (T0 * T1);
    p->ku0 = 1.0 + tmp1;
    T0 = pow(Lnew, m->p.llodvth);
    T1 = pow(W_tmp, m->p.wlodvth);
    tmp1 = m->p.lkvth0 / T0 + m->p.wkvth0 / T1 + m->p.pkvth0 / (T0 * T1);
    p->kvth0 = 1.0 + tmp1;
    p->kvth0 = sqrt(p->kvth0 * p->kvth0 + kDelta);
    T0 = (TRatio - 1.0);
    p->ku0temp = p->ku0 * (1.0 + m->p.tku0 * T0) + kDelta;
    Inv_saref = 1.0 / (m->p.saref + 0.5 * Ldrn);
    Inv_sbref = 1.0 / (m->p.sbref + 0.5 * Ldrn);
    p->inv_od_ref = Inv_saref + Inv_sbref;
    p->rho_ref = m->p.ku0 / p->ku0temp * p->inv_od_ref;
  }
  I->paramPtr = p;

  if ((I->sa > 0.0) && (I->sb > 0.0) && ((nf == 1.0) || ((nf > 1.0) && (I->sd > 0.0)))) {
    int i;
    Inv_sa = 0;
    Inv_sb = 0;
    kvsat = m->p.kvsat;
    if (m->p.kvsat < -1.0) kvsat = -1.0;
    if (m->p.kvsat > 1.0) kvsat = 1.0;
    for (i = 0; i < nf; i++) {
      T0 = 1.0 / nf / (I->sa + 0.5 * Ldrn + i * (I->sd + Ldrn));
      T1 = 1.0 / nf / (I->sb + 0.5 * Ldrn + i * (I->sd + Ldrn));
      Inv_sa += T0;
      Inv_sb += T1;
    }
    Inv_ODeff = Inv_sa + Inv_sb;
    rho = m->p.ku0 / p->ku0temp * Inv_ODeff;
    T0 = (1.0 + rho) / (1.0 + p->rho_ref);
    I->u0temp = p->u0temp * T0;
    T1 = (1.0 + kvsat * rho) / (1.0 + kvsat * p->rho_ref);
    I->vsattemp = p->vsattemp * T1;
    OD_offset = Inv_ODeff - p->inv_od_ref;
    dvth0_lod = m->p.kvth0 / p->kvth0 * OD_offset;
    dk2_lod = m->p.stk2 / pow(p->kvth0, m->p.lodk2) * OD_offset;
    deta0_lod = m->p.steta0 / pow(p->kvth0, m->p.lodeta0) * OD_offset;
    I->vth0 = p->vth0 + dvth0_lod;
    I->eta0 = p->eta0 + deta0_lod;
    I->k2 = p->k2 + dk2_lod;
  } else {
    I->u0temp = p->u0temp;
    I->vth0 = p->vth0;
    I->vsattemp = p->vsattemp;
    I->eta0 = p->eta0;
    I->k2 = p->k2;
  }
  if (m->p.wpemod) {
    if ((!I->sca_given) && (!I->scb_given) && (!I->scc_given)) {
      if ((I->sc_given) && (I->sc > 0.0)) {
        T1 = I->sc + Wdrn;
        T2 = 1.0 / m->p.scref;
        I->sca = m->p.scref * m->p.scref / (I->sc * T1);
        I->scb = ((0.1 * I->sc + 0.01 * m->p.scref) * exp(-10.0 * I->sc * T2)
                  - (0.1 * T1 + 0.01 * m->p.scref) * exp(-10.0 * T1 * T2)) / Wdrn;
        I->scc = ((0.05 * I->sc + 0.0025 * m->p.scref) * exp(-20.0 * I->sc * T2)
                  - (0.05 * T1 + 0.0025 * m->p.scref) * exp(-20.0 * T1 * T2)) / Wdrn;
      }
    }
    sceff = I->sca + m->p.web * I->scb + m->p.wec * I->scc;
    I->vth0 += p->kvth0we * sceff;
    I->k2 += p->k2we * sceff;
    T3 = 1.0 + p->ku0we * sceff;
    if (T3 <= 0.0) T3 = 0.0;
    I->u0temp *= T3;
  }
  I->vth0 += I->delvto;
  I->vfb = p->vfb + m->dtype * I->delvto;
  T3 = m->dtype * I->vth0 - I->vfb - p->phi;
  T4 = T3 + T3;
  T5 = 2.5 * T3;
  I->vtfbphi1 = (m->dtype == 1) ? T4 : T5;
  if (I->vtfbphi1 < 0.0) I->vtfbphi1 = 0.0;
  I->vtfbphi2 = 4.0 * T3;
  if (I->vtfbphi2 < 0.0) I->vtfbphi2 = 0.0;
  if (I->k2 < 0.0) {
    T0 = 0.5 * p->k1 / I->k2;
    I->vbsc = 0.9 * (p->phi - T0 * T0);
    if (I->vbsc > -3.0) I->vbsc = -3.0;
    else if (I->vbsc < -30.0) I->vbsc = -30.0;
  } else {
    I->vbsc = -30.0;
  }
  if (I->vbsc > p->vbm) I->vbsc = p->vbm;
  I->k2ox = I->k2 * toxe / m->p.toxm;
  I->vfbzb = p->vfbzbfactor + m->dtype * I->vth0;
  I->cgso = p->cgso;
  I->cgdo = p->cgdo;
  lnl = log(p->leff * 1.0e6);
  lnw = log(p->weff * 1.0e6);
  lnnf = log(nf);
  bodymode = 5;
  if ((!m->g.rbps0) || (!m->g.rbpd0)) bodymode = 1;
  else if ((!m->g.rbsbx0 && !m->g.rbsby0) || (!m->g.rbdbx0 && !m->g.rbdby0)) bodymode = 3;
  if (I->irbodyMod == 2) {
    if (bodymode == 5) {
      rbsbx = exp(log(m->p.rbsbx0) + m->p.rbsdbxl * lnl + m->p.rbsdbxw * lnw + m->p.rbsdbxnf * lnnf);
      rbsby = exp(log(m->p.rbsby0) + m->p.rbsdbyl * lnl + m->p.rbsdbyw * lnw + m->p.rbsdbynf * lnnf);
      I->rbsb = rbsbx * rbsby / (rbsbx + rbsby);
      rbdbx = exp(log(m->p.rbdbx0) + m->p.rbsdbxl * lnl + m->p.rbsdbxw * lnw + m->p.rbsdbxnf * lnnf);
      rbdby = exp(log(m->p.rbdby0) + m->p.rbsdbyl * lnl + m->p.rbsdbyw * lnw + m->p.rbsdbynf * lnnf);
      I->rbdb = rbdbx * rbdby / (rbdbx + rbdby);
    }
    if ((bodymode == 3) || (bodymode == 5)) {
      I->rbps = exp(log(m->p.rbps0) + m->p.rbpsl * lnl + m->p.rbpsw * lnw + m->p.rbpsnf * lnnf);
      I->rbpd = exp(log(m->p.rbpd0) + m->p.rbpdl * lnl + m->p.rbpdw * lnw + m->p.rbpdnf * lnnf);
    }
    rbpbx = exp(log(m->p.rbpbx0) + m->p.rbpbxl * lnl + m->p.rbpbxw * lnw + m->p.rbpbxnf * lnnf);
    rbpby = exp(log(m->p.rbpby0) + m->p.rbpbyl * lnl + m->p.rbpbyw * lnw + m->p.rbpbynf * lnnf);
    I->rbpb = rbpbx * rbpby / (rbpbx + rbpby);
  }
  if ((I->irbodyMod == 1) || ((I->irbodyMod == 2) && (bodymode == 5))) {
    I->grbdb = (I->rbdb < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbdb;
    I->grbpb = (I->rbpb < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbpb;
    I->grbps = (I->rbps < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbps;
    I->grbsb = (I->rbsb < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbsb;
    I->grbpd = (I->rbpd < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbpd;
  }
  if ((I->irbodyMod == 2) && (bodymode == 3)) {
    I->grbdb = I->grbsb = m->p.gbmin;
    I->grbpb = (I->rbpb < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbpb;
    I->grbps = (I->rbps < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbps;
    I->grbpd = (I->rbpd < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbpd;
  }
  if ((I->irbodyMod == 2) && (bodymode == 1)) {
    I->grbdb = I->grbsb = m->p.gbmin;
    I->grbps = I->grbpd = 1.0e3;
    I->grbpb = (I->rbpb < 1.0e-3) ? 1.0e3 : m->p.gbmin + 1.0 / I->rbpb;
  }
  I->grgeltd = m->p.rshg * (I->xgw + p->weffCJ / 3.0 / I->ngcon) / (I->ngcon * nf * (Lnew - m->p.xgl));
  if (I->grgeltd > 0.0) I->grgeltd = 1.0 / I->grgeltd;
  else I->grgeltd = 1.0e3;
  I->DMCGeff = m->p.dmcg - m->p.dmcgt;
  I->DMCIeff = m->p.dmci;
  I->DMDGeff = m->p.dmdg - m->p.dmcgt;
  if (I->sourcePerimeter_given) {
    if (m->p.perMod == 0) I->Pseff = I->sourcePerimeter;
    else I->Pseff = I->sourcePerimeter - p->weffCJ * nf;
  } else {
    PAeffGeo(nf, I->igeoMod, I->imin, p->weffCJ, I->DMCGeff, I->DMCIeff, I->DMDGeff,
             &I->Pseff, &dumPd, &dumAs, &dumAd);
  }
  if (I->drainPerimeter_given) {
    if (m->p.perMod == 0) I->Pdeff = I->drainPerimeter;
    else I->Pdeff = I->drainPerimeter - p->weffCJ * nf;
  } else {
    PAeffGeo(nf, I->igeoMod, I->imin, p->weffCJ, I->DMCGeff, I->DMCIeff, I->DMDGeff,
             &dumPs, &I->Pdeff, &dumAs, &dumAd);
  }
  if (I->sourceArea_given) I->Aseff = I->sourceArea;
  else PAeffGeo(nf, I->igeoMod, I->imin, p->weffCJ, I->DMCGeff, I->DMCIeff, I->DMDGeff,
                &dumPs, &dumPd, &I->Aseff, &dumAd);
  if (I->drainArea_given) I->Adeff = I->drainArea;
  else PAeffGeo(nf, I->igeoMod, I->imin, p->weffCJ, I->DMCGeff, I->DMCIeff, I->DMDGeff,
                &dumPs, &dumPd, &dumAs, &I->Adeff);
  if (m->p.rdsMod || I->sourceSquares_given) {
    I->sourceConductance = 0.0;
    if (I->sourceSquares_given) I->sourceConductance = m->p.sheetResistance * I->sourceSquares;
    else if (I->irgeoMod > 0)
      RdseffGeo(nf, I->igeoMod, I->irgeoMod, I->imin, p->weffCJ, m->p.sheetResistance,
                I->DMCGeff, I->DMCIeff, I->DMDGeff, 1, &I->sourceConductance);
    else I->sourceConductance = 0.0;
    if (I->sourceConductance > 0.0) I->sourceConductance = 1.0 / I->sourceConductance;
    else I->sourceConductance = 1.0e3;
  } else {
    I->sourceConductance = 0.0;
  }
  if (m->p.rdsMod || I->drainSquares_given) {
    I->drainConductance = 0.0;
    if (I->drainSquares_given) I->drainConductance = m->p.sheetResistance * I->drainSquares;
    else if (I->irgeoMod > 0)
      RdseffGeo(nf, I->igeoMod, I->irgeoMod, I->imin, p->weffCJ, m->p.sheetResistance,
                I->DMCGeff, I->DMCIeff, I->DMDGeff, 0, &I->drainConductance);
    else I->drainConductance = 0.0;
    if (I->drainConductance > 0.0) I->drainConductance = 1.0 / I->drainConductance;
    else I->drainConductance = 1.0e3;
  } else {
    I->drainConductance = 0.0;
  }
  Nvtms = m->vtm * m->p.SjctEmissionCoeff;
  if ((I->Aseff <= 0.0) && (I->Pseff <= 0.0)) SourceSatCurrent = 1.0e-14;
  else SourceSatCurrent = I->Aseff * m->SjctTempSatCurDensity + I->Pseff * m->SjctSidewallTempSatCurDensity
                          + p->weffCJ * nf * m->SjctGateSidewallTempSatCurDensity;
  if (SourceSatCurrent > 0.0) {
    switch ((int)m->p.dioMod) {
      case 0:
        if ((m->p.bvs / Nvtms) > kExpThreshold) I->XExpBVS = m->p.xjbvs * kMinExp;
        else I->XExpBVS = m->p.xjbvs * exp(-m->p.bvs / Nvtms);
        break;
      case 1:
        DioIjthVjmEval(Nvtms, m->p.ijthsfwd, SourceSatCurrent, 0.0, &I->vjsmFwd);
        I->IVjsmFwd = SourceSatCurrent * exp(I->vjsmFwd / Nvtms);
        break;
      case 2:
        if ((m->p.bvs / Nvtms) > kExpThreshold) {
          I->XExpBVS = m->p.xjbvs * kMinExp;
          tmp = kMinExp;
        } else {
          I->XExpBVS = exp(-m->p.bvs / Nvtms);
          tmp = I->XExpBVS;
          I->XExpBVS *= m->p.xjbvs;
        }
        DioIjthVjmEval(Nvtms, m->p.ijthsfwd, SourceSatCurrent, I->XExpBVS, &I->vjsmFwd);
        T0 = exp(I->vjsmFwd / Nvtms);
        I->IVjsmFwd = SourceSatCurrent * (T0 - I->XExpBVS / T0 + I->XExpBVS - 1.0);
        I->SslpFwd = SourceSatCurrent * (T0 + I->XExpBVS / T0) / Nvtms;
        T2 = m->p.ijthsrev / SourceSatCurrent;
        if (T2 < 1.0) T2 = 10.0;
        I->vjsmRev = -m->p.bvs - Nvtms * log((T2 - 1.0) / m->p.xjbvs);
        T1 = m->p.xjbvs * exp(-(m->p.bvs + I->vjsmRev) / Nvtms);
        I->IVjsmRev = SourceSatCurrent * (1.0 + T1);
        I->SslpRev = -SourceSatCurrent * T1 / Nvtms;
        break;
      default: break;
    }
  }
  Nvtmd = m->vtm * m->p.DjctEmissionCoeff;
  if ((I->Adeff <= 0.0) && (I->Pdeff <= 0.0)) DrainSatCurrent = 1.0e-14;
  else DrainSatCurrent = I->Adeff * m->DjctTempSatCurDensity + I->Pdeff * m->DjctSidewallTempSatCurDensity
                         + p->weffCJ * nf * m->DjctGateSidewallTempSatCurDensity;
  if (DrainSatCurrent > 0.0) {
    switch ((int)m->p.dioMod) {
      case 0:
        if ((m->p.bvd / Nvtmd) > kExpThreshold) I->XExpBVD = m->p.xjbvd * kMinExp;
        else I->XExpBVD = m->p.xjbvd * exp(-m->p.bvd / Nvtmd);
        break;
      case 1:
        DioIjthVjmEval(Nvtmd, m->p.ijthdfwd, DrainSatCurrent, 0.0, &I->vjdmFwd);
        I->IVjdmFwd = DrainSatCurrent * exp(I->vjdmFwd / Nvtmd);
        break;
      case 2:
        if ((m->p.bvd / Nvtmd) > kExpThreshold) {
          I->XExpBVD = m->p.xjbvd * kMinExp;
          tmp = kMinExp;
        } else {
          I->XExpBVD = exp(-m->p.bvd / Nvtmd);
          tmp = I->XExpBVD;
          I->XExpBVD *= m->p.xjbvd;
        }
        DioIjthVjmEval(Nvtmd, m->p.ijthdfwd, DrainSatCurrent, I->XExpBVD, &I->vjdmFwd);
        T0 = exp(I->vjdmFwd / Nvtmd);
        I->IVjdmFwd = DrainSatCurrent * (T0 - I->XExpBVD / T0 + I->XExpBVD - 1.0);
        I->DslpFwd = DrainSatCurrent * (T0 + I->XExpBVD / T0) / Nvtmd;
        T2 = m->p.ijthdrev / DrainSatCurrent;
        if (T2 < 1.0) T2 = 10.0;
        I->vjdmRev = -m->p.bvd - Nvtmd * log((T2 - 1.0) / m->p.xjbvd);
        T1 = m->p.xjbvd * exp(-(m->p.bvd + I->vjdmRev) / Nvtmd);
        I->IVjdmRev = DrainSatCurrent * (1.0 + T1);
        I->DslpRev = -DrainSatCurrent * T1 / Nvtmd;
        break;
      default: break;
    }
  }
  T0 = (TRatio - 1.0);
  m->njtsstemp = m->p.njts * (1.0 + m->p.tnjts * T0);
  m->njtsswstemp = m->p.njtssw * (1.0 + m->p.tnjtssw * T0);
  m->njtsswgstemp = m->p.njtsswg * (1.0 + m->p.tnjtsswg * T0);
  m->njtsdtemp = m->p.njtsd * (1.0 + m->p.tnjtsd * T0);
  m->njtsswdtemp = m->p.njtsswd * (1.0 + m->p.tnjtsswd * T0);
  m->njtsswgdtemp = m->p.njtsswgd * (1.0 + m->p.tnjtsswgd * T0);
  T7 = Eg0 / m->vtm * T0;
#define XA_DEXP2(A, B) do { if ((A) > kExpThreshold) (B) = kMaxExp * (1.0 + (A) - kExpThreshold); \
    else if ((A) < -kExpThreshold) (B) = kMinExp; else (B) = exp(A); } while (0)
  T9 = m->p.xtss * T7;
  XA_DEXP2(T9, T1);
  T9 = m->p.xtsd * T7;
  XA_DEXP2(T9, T2);
  T9 = m->p.xtssws * T7;
  XA_DEXP2(T9, T3);
  T9 = m->p.xtsswd * T7;
  XA_DEXP2(T9, T4);
  T9 = m->p.xtsswgs * T7;
  XA_DEXP2(T9, T5);
  T9 = m->p.xtsswgd * T7;
  XA_DEXP2(T9, T6);
  T10 = p->weffCJ * nf;
  I->SjctTempRevSatCur = T1 * I->Aseff * m->p.jtss;
  I->DjctTempRevSatCur = T2 * I->Adeff * m->p.jtsd;
  I->SswTempRevSatCur = T3 * I->Pseff * m->p.jtssws;
  I->DswTempRevSatCur = T4 * I->Pdeff * m->p.jtsswd;
  I->SswgTempRevSatCur = T5 * T10 * m->p.jtsswgs;
  I->DswgTempRevSatCur = T6 * T10 * m->p.jtsswgd;
  if (m->p.mtrlMod) {
    /* TOXP from EOT (:7614-7716): Vgs_eff at VDDEOT with poly depletion,
     * Vth at Vds = Vbs = 0, then the charge centroid iterated */
    double vddeot, Vgs_eff, V0, lt1, ltw, Theta0, Delt_vth, TempRatio, Vth_NarrowW, Lpe_Vb, Vth;
    double n, tmp3, Vgsteff, toxpf, toxpi, Tcen;
    int niter;
    tmp2 = I->vfb + p->phi;
    vddeot = m->dtype * m->p.vddeot;
    T0 = m->p.epsrgate * kEps0;
    if ((p->ngate > 1.0e18) && (p->ngate < 1.0e25) && (vddeot > tmp2) && (T0 != 0)) {
      T1 = 1.0e6 * kQ * T0 * p->ngate / (m->coxe * m->coxe);
      T8 = vddeot - tmp2;
      T4 = sqrt(1.0 + 2.0 * T8 / T1);
      T2 = 2.0 * T8 / (T4 + 1.0);
      T3 = 0.5 * T2 * T2 / T1;
      T7 = 1.12 - T3 - 0.05;
      T6 = sqrt(T7 * T7 + 0.224);
      T5 = 1.12 - 0.5 * (T7 + T6);
      Vgs_eff = vddeot - T5;
    } else {
      Vgs_eff = vddeot;
    }
    V0 = p->vbi - p->phi;
    lt1 = m->factor1 * p->sqrtXdep0;
    ltw = lt1;
    T0 = p->dvt1 * p->leff / lt1;
    if (T0 < kExpThreshold) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * kMinExp;
      Theta0 = T1 / T4;
    } else {
      Theta0 = 1.0 / (kMaxExp - 2.0);
    }
    Delt_vth = p->dvt0 * Theta0 * V0;
    T0 = p->dvt1w * p->weff * p->leff / ltw;
    if (T0 < kExpThreshold) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * kMinExp;
      T5 = T1 / T4;
    } else {
      T5 = 1.0 / (kMaxExp - 2.0);
    }
    T2 = p->dvt0w * T5 * V0;
    TempRatio = I->temp / m->p.tnom - 1.0;
    T0 = sqrt(1.0 + p->lpe0 / p->leff);
    T1 = p->k1ox * (T0 - 1.0) * p->sqrtPhi + (p->kt1 + p->kt1l / p->leff) * TempRatio;
    Vth_NarrowW = toxe * p->phi / (p->weff + p->w0);
    Lpe_Vb = sqrt(1.0 + p->lpeb / p->leff);
    Vth = m->dtype * I->vth0 + (p->k1ox - p->k1) * p->sqrtPhi * Lpe_Vb - Delt_vth - T2
          + p->k3 * Vth_NarrowW + T1;
    tmp1 = epssub / p->Xdep0;
    tmp2 = p->nfactor * tmp1;
    tmp3 = (tmp2 + p->cdsc * Theta0 + p->cit) / m->coxe;
    if (tmp3 >= -0.5) {
      n = 1.0 + tmp3;
    } else {
      T0 = 1.0 / (3.0 + 8.0 * tmp3);
      n = (1.0 + 3.0 * tmp3) * T0;
    }
    if (p->dvtp0 > 0.0) {
      T3 = p->leff + p->dvtp0 * 2.0;
      if (m->p.tempMod < 2) T4 = m->vtm * log(p->leff / T3);
      else T4 = m->vtm0 * log(p->leff / T3);
      Vth -= n * T4;
    }
    Vgsteff = Vgs_eff - Vth;
    niter = 0;
    toxpf = toxe;
    do {
      toxpi = toxpf;
      tmp2 = 2.0e8 * toxpf;
      T0 = (Vgsteff + I->vtfbphi2) / tmp2;
      T1 = 1.0 + exp(m->p.bdos * 0.7 * log(T0));
      Tcen = m->p.ados * 1.9e-9 / T1;
      toxpf = toxe - epsrox / m->p.epsrsub * Tcen;
      niter++;
    } while ((niter <= 4) && (fabs(toxpf - toxpi) > 1e-12));
    m->p.toxp = toxpf;
    m->coxp = epsrox * kEps0 / m->p.toxp;
  }
}

/* Instance ctor defaults (N_DEV_MOSFET_B4.C:4585-4618) and
 * Instance::processParams (:3621-3794).  Slots of g.ival follow the B4I
 * order of the parameter table. */
int instance_init(Group& g, Inst& iref, int64_t i) {
  Inst* I = &iref;
  Model* m = I->model;
  int k = 0;
  double Rtot;
  int createNode;
#define B4I(NAME, mem, def) I->mem = g.ival[k][i]; I->mem##_given = g.igiven[k]; ++k;
#include "xa_bsim4_params.def"
#undef B4I
  if (!I->rbdb_given) I->rbdb = m->p.rbdb;
  if (!I->rbsb_given) I->rbsb = m->p.rbsb;
  if (!I->rbpb_given) I->rbpb = m->p.rbpb;
  if (!I->rbps_given) I->rbps = m->p.rbps;
  if (!I->rbpd_given) I->rbpd = m->p.rbpd;
  if (!I->xgw_given) I->xgw = m->p.xgw;
  if (!I->ngcon_given) I->ngcon = m->p.ngcon;
  if (!I->sd_given) I->sd = 2.0 * m->p.dmcg;
  if (!I->temp_given) I->temp = g.opt.temp;
  if (!I->drainArea_given) I->drainArea = g.opt.defad;
  if (!I->sourceArea_given) I->sourceArea = g.opt.defas;
  I->irbodyMod = (int)I->rbodyMod;
  I->irgateMod = (int)I->rgateMod;
  I->igeoMod = (int)I->geoMod;
  I->irgeoMod = (int)I->rgeoMod;
  I->itrnqsMod = (int)I->trnqsMod;
  I->iacnqsMod = (int)I->acnqsMod;
  I->imin = (int)I->min;
  if (!I->rbodyMod_given) I->irbodyMod = (int)m->p.rbodyMod;
  else if (I->irbodyMod != 0 && I->irbodyMod != 1 && I->irbodyMod != 2) I->irbodyMod = (int)m->p.rbodyMod;
  if (!I->rgateMod_given) I->irgateMod = (int)m->p.rgateMod;
  else if (I->irgateMod < 0 || I->irgateMod > 3) I->irgateMod = (int)m->p.rgateMod;
  if (!I->geoMod_given) I->igeoMod = (int)m->p.geoMod;
  if (!I->rgeoMod_given) I->irgeoMod = 0;
  if (!I->trnqsMod_given) I->itrnqsMod = (int)m->p.trnqsMod;
  else if (I->itrnqsMod != 0 && I->itrnqsMod != 1) I->itrnqsMod = (int)m->p.trnqsMod;
  if (!I->acnqsMod_given) I->iacnqsMod = (int)m->p.acnqsMod;
  else if (I->iacnqsMod != 0 && I->iacnqsMod != 1) I->iacnqsMod = (int)m->p.acnqsMod;
  update_temperature(I, I->temp, g.opt);
  createNode = 0;
  if (m->p.rdsMod != 0) createNode = 1;
  else if (m->p.sheetResistance > 0) {
    if (I->drainSquares_given && I->drainSquares > 0) createNode = 1;
    else if (!I->drainSquares_given && (I->irgeoMod != 0)) {
      RdseffGeo(I->nf, I->igeoMod, I->irgeoMod, I->imin, I->w, m->p.sheetResistance, I->DMCGeff,
                I->DMCIeff, I->DMDGeff, 0, &Rtot);
      if (Rtot > 0) createNode = 1;
    }
  }
  I->drainMOSFET_B4Exists = createNode != 0;
  createNode = 0;
  if (m->p.rdsMod != 0) createNode = 1;
  else if (m->p.sheetResistance > 0) {
    if (I->sourceSquares_given && I->sourceSquares > 0) createNode = 1;
    else if (!I->sourceSquares_given && (I->irgeoMod != 0)) {
      RdseffGeo(I->nf, I->igeoMod, I->irgeoMod, I->imin, I->w, m->p.sheetResistance, I->DMCGeff,
                I->DMCIeff, I->DMDGeff, 1, &Rtot);
      if (Rtot > 0) createNode = 1;
    }
  }
  I->sourceMOSFET_B4Exists = createNode != 0;
  /* numIntVars / numStateVars (:3754-3788) */
  g.n_ext[i] = 4;
  g.n_int[i] = (I->drainMOSFET_B4Exists ? 1 : 0) + (I->sourceMOSFET_B4Exists ? 1 : 0)
               + ((I->irgateMod == 1 || I->irgateMod == 2) ? 1 : (I->irgateMod == 3) ? 2 : 0)
               + (I->itrnqsMod ? 1 : 0) + (I->irbodyMod ? 3 : 0)
               + (I->icVBS_given ? 1 : 0) + (I->icVDS_given ? 1 : 0) + (I->icVGS_given ? 1 : 0);
  g.n_sta[i] = 3 + (I->irgateMod == 3 ? 1 : 0) + (I->irbodyMod ? 2 : 0) + (I->itrnqsMod ? 2 : 0);
  g.n_sto[i] = 14;   /* the BSIM4 Master loads write no lead currents */
  return XA_OK;
}

}  // namespace bsim4
}  // namespace xa
