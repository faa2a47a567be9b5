/* or_bsim4_setup.c -- oracle restatement of BSIM4 model/instance setup
 * (test infrastructure only).
 * Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET_B4.C
 *   Model ctor defaults :15661-16056, Model::processParams :14740-14780,
 *   Instance ctor defaults :4585-4618, Instance::processParams :3621-3794,
 *   updateTemperature :5640-7734, polyDepletion-free helpers
 *   DioIjthVjmEval :14292-14303, NumFingerDiff :14322-14351,
 *   PAeffGeo :14364-14462, RdseffGeo :14473-14588, RdsEndIso :14599-14657,
 *   RdsEndSha :14668-14726. */
#include <stddef.h>
#include <stdio.h>
#include "or_bsim4.h"

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
    *nuIntD = *nuIntS = 2.0 * or_max((nf_arg - 1.0) / 2.0, 0.0);
  } else {
    if (minSD == 1) {
      *nuEndD = 2.0;
      *nuIntD = 2.0 * or_max((nf_arg / 2.0 - 1.0), 0.0);
      *nuEndS = 0.0;
      *nuIntS = nf_arg;
    } else {
      *nuEndD = 0.0;
      *nuIntD = nf_arg;
      *nuEndS = 2.0;
      *nuIntS = 2.0 * or_max((nf_arg / 2.0 - 1.0), 0.0);
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
int b4_model_init(b4model* m, const or_card* c, const xa_device_options* opt) {
  int k = 0;
  memset(m, 0, sizeof(*m));
#define B4M(NAME, mem, def) m->p.mem = c->v[k]; m->g.mem = c->given[k]; ++k;
#include "../include/xa_bsim4_params.def"
#undef B4M
  /* pseudo-parameter TYPE follows the table: the .model nmos/pmos keyword */
  m->dtype = (c->v[k] < 0) ? -1 : 1;
  if (!MG(tnom)) M(tnom) = opt->tnom;
  if (!MG(toxp)) M(toxp) = M(toxe);
  if (!MG(toxm)) M(toxm) = M(toxe);
  if (!MG(dsub)) M(dsub) = M(drout);
  if (!MG(vth0)) M(vth0) = (m->dtype == 1) ? 0.7 : -0.7;
  if (!MG(vddeot)) M(vddeot) = (m->dtype == 1) ? 1.5 : -1.5;
  if (!MG(eu)) M(eu) = (m->dtype == 1) ? 1.67 : 1.0;
  if (!MG(ua)) M(ua) = (M(mobMod) == 2) ? 1.0E-15 : 1.0E-9;
  if (!MG(uc)) M(uc) = (M(mobMod) == 1) ? -0.0465 : -0.0465E-9;
  if (!MG(uc1)) M(uc1) = (M(mobMod) == 1) ? -0.056 : -0.056E-9;
  if (!MG(u0)) M(u0) = (m->dtype == 1) ? 0.067 : 0.025;
  if (!MG(aigc)) M(aigc) = (m->dtype == 1) ? 1.36E-2 : 9.80E-3;
  if (!MG(bigc)) M(bigc) = (m->dtype == 1) ? 1.71E-3 : 7.59E-4;
  if (!MG(cigc)) M(cigc) = (m->dtype == 1) ? 0.075 : 0.03;
  if (MG(aigsd)) {
    M(aigs) = M(aigd) = M(aigsd);
  } else {
    M(aigsd) = (m->dtype == 1) ? 1.36E-2 : 9.80E-3;
    if (!MG(aigs)) M(aigs) = M(aigsd);
    if (!MG(aigd)) M(aigd) = M(aigsd);
  }
  if (MG(bigsd)) {
    M(bigs) = M(bigd) = M(bigsd);
  } else {
    M(bigsd) = (m->dtype == 1) ? 1.71E-3 : 7.59E-4;
    if (!MG(bigs)) M(bigs) = M(bigsd);
    if (!MG(bigd)) M(bigd) = M(bigsd);
  }
  if (MG(cigsd)) {
    M(cigs) = M(cigd) = M(cigsd);
  } else {
    M(cigsd) = (m->dtype == 1) ? 0.075 : 0.03;
    if (!MG(cigs)) M(cigs) = M(cigsd);
    if (!MG(cigd)) M(cigd) = M(cigsd);
  }
  if (!MG(ijthdfwd)) M(ijthdfwd) = M(ijthsfwd);
  if (!MG(ijthdrev)) M(ijthdrev) = M(ijthsrev);
  if (!MG(xjbvd)) M(xjbvd) = M(xjbvs);
  if (!MG(bvd)) M(bvd) = M(bvs);
  if (!MG(ckappad)) M(ckappad) = M(ckappas);
  if (!MG(dmci)) M(dmci) = M(dmcg);
  if (!MG(Llc)) M(Llc) = M(Ll);
  if (!MG(Lwc)) M(Lwc) = M(Lw);
  if (!MG(Lwlc)) M(Lwlc) = M(Lwl);
  if (!MG(Wlc)) M(Wlc) = M(Wl);
  if (!MG(Wwc)) M(Wwc) = M(Ww);
  if (!MG(Wwlc)) M(Wwlc) = M(Wwl);
  if (!MG(dwc)) M(dwc) = M(Wint);
  if (!MG(dlc)) M(dlc) = M(Lint);
  if (!MG(agisl)) { if (MG(agidl)) M(agisl) = M(agidl); }
  if (!MG(bgisl)) { if (MG(bgidl)) M(bgisl) = M(bgidl); }
  if (!MG(cgisl)) { if (MG(cgidl)) M(cgisl) = M(cgidl); }
  if (!MG(egisl)) { if (MG(egidl)) M(egisl) = M(egidl); }
  if (!MG(dlcig)) M(dlcig) = M(Lint);
  if (!MG(dlcigd)) {
    if (!MG(dlcig)) M(dlcigd) = M(Lint);
    else M(dlcigd) = M(dlcig);
  }
  if (!MG(dwj)) M(dwj) = M(dwc);
  if (!MG(cf)) M(cf) = 2.0 * M(epsrox) * (B4_EPS0 / 3.14159265358979323846) * log(1.0 + 0.4E-6 / M(toxe));
  if (!MG(DjctSatCurDensity)) M(DjctSatCurDensity) = M(SjctSatCurDensity);
  if (!MG(DjctSidewallSatCurDensity)) M(DjctSidewallSatCurDensity) = M(SjctSidewallSatCurDensity);
  if (!MG(DjctGateSidewallSatCurDensity)) M(DjctGateSidewallSatCurDensity) = M(SjctGateSidewallSatCurDensity);
  if (!MG(DbulkJctPotential)) M(DbulkJctPotential) = M(SbulkJctPotential);
  if (!MG(DjctEmissionCoeff)) M(DjctEmissionCoeff) = M(SjctEmissionCoeff);
  if (!MG(DjctTempExponent)) M(DjctTempExponent) = M(SjctTempExponent);
  if (!MG(DbulkJctBotGradingCoeff)) M(DbulkJctBotGradingCoeff) = M(SbulkJctBotGradingCoeff);
  if (!MG(DbulkJctSideGradingCoeff)) M(DbulkJctSideGradingCoeff) = M(SbulkJctSideGradingCoeff);
  if (!MG(SbulkJctGateSideGradingCoeff)) M(SbulkJctGateSideGradingCoeff) = M(SbulkJctSideGradingCoeff);
  if (!MG(DbulkJctGateSideGradingCoeff)) M(DbulkJctGateSideGradingCoeff) = M(SbulkJctGateSideGradingCoeff);
  if (!MG(DsidewallJctPotential)) M(DsidewallJctPotential) = M(SsidewallJctPotential);
  if (!MG(SGatesidewallJctPotential)) M(SGatesidewallJctPotential) = M(SsidewallJctPotential);
  if (!MG(DGatesidewallJctPotential)) M(DGatesidewallJctPotential) = M(DsidewallJctPotential);
  if (!MG(DunitAreaJctCap)) M(DunitAreaJctCap) = M(SunitAreaJctCap);
  if (!MG(DunitLengthSidewallJctCap)) M(DunitLengthSidewallJctCap) = M(SunitLengthSidewallJctCap);
  if (!MG(SunitLengthGateSidewallJctCap)) M(SunitLengthGateSidewallJctCap) = M(SunitLengthSidewallJctCap);
  if (!MG(DunitLengthGateSidewallJctCap)) M(DunitLengthGateSidewallJctCap) = M(SunitLengthGateSidewallJctCap);
  if (!MG(jtsd)) M(jtsd) = M(jtss);
  if (!MG(jtsswd)) M(jtsswd) = M(jtssws);
  if (!MG(jtsswgd)) M(jtsswgd) = M(jtsswgs);
  if (!MG(njtsd)) { if (MG(njts)) M(njtsd) = M(njts); }
  if (!MG(njtsswd)) { if (MG(njtssw)) M(njtsswd) = M(njtssw); }
  if (!MG(njtsswgd)) { if (MG(njtsswg)) M(njtsswgd) = M(njtsswg); }
  if (!MG(xtsd)) M(xtsd) = M(xtss);
  if (!MG(xtsswd)) M(xtsswd) = M(xtssws);
  if (!MG(xtsswgd)) M(xtsswgd) = M(xtsswgs);
  if (!MG(tnjtsd)) { if (MG(tnjts)) M(tnjtsd) = M(tnjts); }
  if (!MG(tnjtsswd)) { if (MG(tnjtssw)) M(tnjtsswd) = M(tnjtssw); }
  if (!MG(tnjtsswgd)) { if (MG(tnjtsswg)) M(tnjtsswgd) = M(tnjtsswg); }
  if (!MG(vtsd)) M(vtsd) = M(vtss);
  if (!MG(vtsswd)) M(vtsswd) = M(vtssws);
  if (!MG(vtsswgd)) M(vtsswgd) = M(vtsswgs);
  if (!MG(lagisl)) { if (MG(lagidl)) M(lagisl) = M(lagidl); }
  if (!MG(lbgisl)) { if (MG(lbgidl)) M(lbgisl) = M(lbgidl); }
  if (!MG(lcgisl)) { if (MG(lcgidl)) M(lcgisl) = M(lcgidl); }
  if (!MG(legisl)) { if (MG(legidl)) M(legisl) = M(legidl); }
  if (!(!MG(aigsd) && (MG(aigs) || MG(aigd)))) M(laigs) = M(laigd) = M(laigsd);
  if (!(!MG(bigsd) && (MG(bigs) || MG(bigd)))) M(lbigs) = M(lbigd) = M(lbigsd);
  if (!(!MG(cigsd) && (MG(cigs) || MG(cigd)))) M(lcigs) = M(lcigd) = M(lcigsd);
  if (!MG(wagisl)) { if (MG(wagidl)) M(wagisl) = M(wagidl); }
  if (!MG(wbgisl)) { if (MG(wbgidl)) M(wbgisl) = M(wbgidl); }
  if (!MG(wcgisl)) { if (MG(wcgidl)) M(wcgisl) = M(wcgidl); }
  if (!MG(wegisl)) { if (MG(wegidl)) M(wegisl) = M(wegidl); }
  if (!(!MG(aigsd) && (MG(aigs) || MG(aigd)))) M(waigs) = M(waigd) = M(waigsd);
  if (!(!MG(bigsd) && (MG(bigs) || MG(bigd)))) M(wbigs) = M(wbigd) = M(wbigsd);
  if (!(!MG(cigsd) && (MG(cigs) || MG(cigd)))) M(wcigs) = M(wcigd) = M(wcigsd);
  if (!MG(pagisl)) { if (MG(pagidl)) M(pagisl) = M(pagidl); }
  if (!MG(pbgisl)) { if (MG(pbgidl)) M(pbgisl) = M(pbgidl); }
  if (!MG(pcgisl)) { if (MG(pcgidl)) M(pcgisl) = M(pcgidl); }
  if (!MG(pegisl)) { if (MG(pegidl)) M(pegisl) = M(pegidl); }
  if (!(!MG(aigsd) && (MG(aigs) || MG(aigd)))) M(paigs) = M(paigd) = M(paigsd);
  if (!(!MG(bigsd) && (MG(bigs) || MG(bigd)))) M(pbigs) = M(pbigd) = M(pbigsd);
  if (!(!MG(cigsd) && (MG(cigs) || MG(cigd)))) M(pcigs) = M(pcigd) = M(pcigsd);
  /* Model::processParams */
  if (M(SbulkJctPotential) < 0.1) M(SbulkJctPotential) = 0.1;
  if (M(SsidewallJctPotential) < 0.1) M(SsidewallJctPotential) = 0.1;
  if (M(SGatesidewallJctPotential) < 0.1) M(SGatesidewallJctPotential) = 0.1;
  if (M(DbulkJctPotential) < 0.1) M(DbulkJctPotential) = 0.1;
  if (M(DsidewallJctPotential) < 0.1) M(DsidewallJctPotential) = 0.1;
  if (M(DGatesidewallJctPotential) < 0.1) M(DGatesidewallJctPotential) = 0.1;
  /* option subset implemented by this restatement */
  /* updateTemperature rejects other junction models (N_DEV_MOSFET_B4.C:7515) */
  if (M(dioMod) != 0 && M(dioMod) != 1 && M(dioMod) != 2)
    return or_fail(XA_ERR_ARG, "bsim4: specified dioMod not matched");
  /* the evaluation has branches for these values only (any other would
   * leave its charges or gate-current terms unset) */
  if (M(capMod) != 0 && M(capMod) != 1 && M(capMod) != 2)
    return or_fail(XA_ERR_UNSUPPORTED, "bsim4: capMod must be 0, 1 or 2");
  if (M(igcMod) != 0 && M(igcMod) != 1 && M(igcMod) != 2)
    return or_fail(XA_ERR_UNSUPPORTED, "bsim4: igcMod must be 0, 1 or 2");
  if (M(igbMod) != 0 && M(igbMod) != 1) return or_fail(XA_ERR_UNSUPPORTED, "bsim4: igbMod must be 0 or 1");
  return XA_OK;
}

static b4size* size_lookup(b4model* m, double l, double w, double nf) {
  int k;
  b4size* hit = NULL;
  for (k = 0; k < m->nsizes; ++k)   /* last match wins, as the reference's scan */
    if (m->sizes[k]->Length == l && m->sizes[k]->Width == w && m->sizes[k]->NFinger == nf)
      hit = m->sizes[k];
  return hit;
}

/* Instance::updateTemperature (N_DEV_MOSFET_B4.C:5640-7734) */
void b4_update_temperature(b4inst* I, double temp_tmp, const xa_device_options* opt) {
  b4model* m = I->model;
  b4size* p;
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
  Tnom = M(tnom);
  TRatio = I->temp / Tnom;
  if (M(mtrlMod) == 0) {
    if ((MG(toxe)) && (MG(toxp)) && (MG(dtox)) && (M(toxe) != (M(toxp) + M(dtox)))) {
    } else if ((MG(toxe)) && (!MG(toxp))) {
      M(toxp) = M(toxe) - M(dtox);
    } else if ((!MG(toxe)) && (MG(toxp))) {
      M(toxe) = M(toxp) + M(dtox);
    }
  }
  if (M(mtrlMod)) {
    epsrox = 3.9;
    toxe = M(eot);
    epssub = B4_EPS0 * M(epsrsub);
  } else {
    epsrox = M(epsrox);
    toxe = M(toxe);
    epssub = B4_EPSSI;
  }
  m->coxe = epsrox * B4_EPS0 / toxe;
  if (M(mtrlMod) == 0) m->coxp = M(epsrox) * B4_EPS0 / M(toxp);
  if (!MG(cgdo)) {
    if (MG(dlc) && (M(dlc) > 0.0)) M(cgdo) = M(dlc) * m->coxe - M(cgdl);
    else M(cgdo) = 0.6 * M(xj) * m->coxe;
  }
  if (!MG(cgso)) {
    if (MG(dlc) && (M(dlc) > 0.0)) M(cgso) = M(dlc) * m->coxe - M(cgsl);
    else M(cgso) = 0.6 * M(xj) * m->coxe;
  }
  if (!MG(cgbo)) M(cgbo) = 2.0 * M(dwc) * m->coxe;
  m->vcrit = B4_vt0 * log(B4_vt0 / (sqrt(2.0) * 1.0e-14));
  m->factor1 = sqrt(epssub / (epsrox * B4_EPS0) * toxe);
  Vtm0 = m->vtm0 = B4_KboQ * Tnom;
  if (M(mtrlMod) == 0) {
    Eg0 = 1.16 - 7.02e-4 * Tnom * Tnom / (Tnom + 1108.0);
    ni = 1.45e10 * (Tnom / 300.15) * sqrt(Tnom / 300.15) * exp(21.5565981 - Eg0 / (2.0 * Vtm0));
  } else {
    Eg0 = M(bg0sub) - M(tbgasub) * Tnom * Tnom / (Tnom + M(tbgbsub));
    T0 = M(bg0sub) - M(tbgasub) * 90090.0225 / (300.15 + M(tbgbsub));
    ni = M(ni0sub) * (Tnom / 300.15) * sqrt(Tnom / 300.15) * exp((T0 - Eg0) / (2.0 * Vtm0));
  }
  m->Eg0 = Eg0;
  m->vtm = B4_KboQ * I->temp;
  if (M(mtrlMod) == 0) Eg = 1.16 - 7.02e-4 * I->temp * I->temp / (I->temp + 1108.0);
  else Eg = M(bg0sub) - M(tbgasub) * I->temp * I->temp / (I->temp + M(tbgbsub));
  if (I->temp != Tnom) {
    T0 = Eg0 / Vtm0 - Eg / m->vtm;
    T1 = log(I->temp / Tnom);
    T2 = T0 + M(SjctTempExponent) * T1;
    T3 = exp(T2 / M(SjctEmissionCoeff));
    m->SjctTempSatCurDensity = M(SjctSatCurDensity) * T3;
    m->SjctSidewallTempSatCurDensity = M(SjctSidewallSatCurDensity) * T3;
    m->SjctGateSidewallTempSatCurDensity = M(SjctGateSidewallSatCurDensity) * T3;
    T2 = T0 + M(DjctTempExponent) * T1;
    T3 = exp(T2 / M(DjctEmissionCoeff));
    m->DjctTempSatCurDensity = M(DjctSatCurDensity) * T3;
    m->DjctSidewallTempSatCurDensity = M(DjctSidewallSatCurDensity) * T3;
    m->DjctGateSidewallTempSatCurDensity = M(DjctGateSidewallSatCurDensity) * T3;
  } else {
    m->SjctTempSatCurDensity = M(SjctSatCurDensity);
    m->SjctSidewallTempSatCurDensity = M(SjctSidewallSatCurDensity);
    m->SjctGateSidewallTempSatCurDensity = M(SjctGateSidewallSatCurDensity);
    m->DjctTempSatCurDensity = M(DjctSatCurDensity);
    m->DjctSidewallTempSatCurDensity = M(DjctSidewallSatCurDensity);
    m->DjctGateSidewallTempSatCurDensity = M(DjctGateSidewallSatCurDensity);
  }
  if (m->SjctTempSatCurDensity < 0.0) m->SjctTempSatCurDensity = 0.0;
  if (m->SjctSidewallTempSatCurDensity < 0.0) m->SjctSidewallTempSatCurDensity = 0.0;
  if (m->SjctGateSidewallTempSatCurDensity < 0.0) m->SjctGateSidewallTempSatCurDensity = 0.0;
  if (m->DjctTempSatCurDensity < 0.0) m->DjctTempSatCurDensity = 0.0;
  if (m->DjctSidewallTempSatCurDensity < 0.0) m->DjctSidewallTempSatCurDensity = 0.0;
  if (m->DjctGateSidewallTempSatCurDensity < 0.0) m->DjctGateSidewallTempSatCurDensity = 0.0;
  delTemp = I->temp - Tnom;
  T0 = M(tcj) * delTemp;
  if (T0 >= -1.0) {
    m->SunitAreaTempJctCap = M(SunitAreaJctCap) * (1.0 + T0);
    m->DunitAreaTempJctCap = M(DunitAreaJctCap) * (1.0 + T0);
  } else {
    if (M(SunitAreaJctCap) > 0.0) m->SunitAreaTempJctCap = 0.0;
    if (M(DunitAreaJctCap) > 0.0) m->DunitAreaTempJctCap = 0.0;
  }
  T0 = M(tcjsw) * delTemp;
  if (T0 >= -1.0) {
    m->SunitLengthSidewallTempJctCap = M(SunitLengthSidewallJctCap) * (1.0 + T0);
    m->DunitLengthSidewallTempJctCap = M(DunitLengthSidewallJctCap) * (1.0 + T0);
  } else {
    if (M(SunitLengthSidewallJctCap) > 0.0) m->SunitLengthSidewallTempJctCap = 0.0;
    if (M(DunitLengthSidewallJctCap) > 0.0) m->DunitLengthSidewallTempJctCap = 0.0;
  }
  T0 = M(tcjswg) * delTemp;
  if (T0 >= -1.0) {
    m->SunitLengthGateSidewallTempJctCap = M(SunitLengthGateSidewallJctCap) * (1.0 + T0);
    m->DunitLengthGateSidewallTempJctCap = M(DunitLengthGateSidewallJctCap) * (1.0 + T0);
  } else {
    if (M(SunitLengthGateSidewallJctCap) > 0.0) m->SunitLengthGateSidewallTempJctCap = 0.0;
    if (M(DunitLengthGateSidewallJctCap) > 0.0) m->DunitLengthGateSidewallTempJctCap = 0.0;
  }
  m->PhiBS = M(SbulkJctPotential) - M(tpb) * delTemp;
  if (m->PhiBS < 0.01) m->PhiBS = 0.01;
  m->PhiBD = M(DbulkJctPotential) - M(tpb) * delTemp;
  if (m->PhiBD < 0.01) m->PhiBD = 0.01;
  m->PhiBSWS = M(SsidewallJctPotential) - M(tpbsw) * delTemp;
  if (m->PhiBSWS <= 0.01) m->PhiBSWS = 0.01;
  m->PhiBSWD = M(DsidewallJctPotential) - M(tpbsw) * delTemp;
  if (m->PhiBSWD <= 0.01) m->PhiBSWD = 0.01;
  m->PhiBSWGS = M(SGatesidewallJctPotential) - M(tpbswg) * delTemp;
  if (m->PhiBSWGS <= 0.01) m->PhiBSWGS = 0.01;
  m->PhiBSWGD = M(DGatesidewallJctPotential) - M(tpbswg) * delTemp;
  if (m->PhiBSWGD <= 0.01) m->PhiBSWGD = 0.01;
  if (M(ijthdfwd) <= 0.0) M(ijthdfwd) = 0.1;
  if (M(ijthsfwd) <= 0.0) M(ijthsfwd) = 0.1;
  if (M(ijthdrev) <= 0.0) M(ijthdrev) = 0.1;
  if (M(ijthsrev) <= 0.0) M(ijthsrev) = 0.1;
  if ((M(xjbvd) <= 0.0) && (M(dioMod) == 2)) M(xjbvd) = 1.0;
  else if ((M(xjbvd) < 0.0) && (M(dioMod) == 0)) M(xjbvd) = 1.0;
  if (M(bvd) <= 0.0) M(bvd) = 10.0;
  if ((M(xjbvs) <= 0.0) && (M(dioMod) == 2)) M(xjbvs) = 1.0;
  else if ((M(xjbvs) < 0.0) && (M(dioMod) == 0)) M(xjbvs) = 1.0;
  if (M(bvs) <= 0.0) M(bvs) = 10.0;

  Ldrn = I->l;
  Wdrn = I->w / nf;
  p = size_lookup(m, I->l, I->w, nf);
  Lnew = I->l + M(xl);
  Wnew = I->w / nf + M(xw);
  if (p == NULL) {
    p = (b4size*)calloc(1, sizeof(b4size));
    if (m->nsizes == m->capsizes) {
      m->capsizes = m->capsizes ? 2 * m->capsizes : 16;
      m->sizes = (b4size**)realloc(m->sizes, sizeof(b4size*) * m->capsizes);
    }
    m->sizes[m->nsizes++] = p;
    p->referenceTemperature = temp_tmp;
    p->Length = I->l;
    p->Width = I->w;
    p->NFinger = nf;
    T0 = pow(Lnew, M(Lln));
    T1 = pow(Wnew, M(Lwn));
    tmp1 = M(Ll) / T0 + M(Lw) / T1 + M(Lwl) / (T0 * T1);
    p->dl = M(Lint) + tmp1;
    tmp2 = M(Llc) / T0 + M(Lwc) / T1 + M(Lwlc) / (T0 * T1);
    p->dlc = M(dlc) + tmp2;
    T2 = pow(Lnew, M(Wln));
    T3 = pow(Wnew, M(Wwn));
    tmp1 = M(Wl) / T2 + M(Ww) / T3 + M(Wwl) / (T2 * T3);
    p->dw = M(Wint) + tmp1;
    tmp2 = M(Wlc) / T2 + M(Wwc) / T3 + M(Wwlc) / (T2 * T3);
    p->dwc = M(dwc) + tmp2;
    p->dwj = M(dwj) + tmp2;
    p->leff = Lnew - 2.0 * p->dl;
    p->weff = Wnew - 2.0 * p->dw;
    p->leffCV = Lnew - 2.0 * p->dlc;
    p->weffCV = Wnew - 2.0 * p->dwc;
    p->weffCJ = Wnew - 2.0 * p->dwj;
    if (M(binUnit) == 1) {
      Inv_L = 1.0e-6 / p->leff;
      Inv_W = 1.0e-6 / p->weff;
      Inv_LW = 1.0e-12 / (p->leff * p->weff);
    } else {
      Inv_L = 1.0 / p->leff;
      Inv_W = 1.0 / p->weff;
      Inv_LW = 1.0 / (p->leff * p->weff);
    }
#define B4BIN(sm, base, l, w, pp) p->sm = M(base) + M(l) * Inv_L + M(w) * Inv_W + M(pp) * Inv_LW;
#include "../include/xa_bsim4_binning.def"
#undef B4BIN
    p->abulkCVfactor = 1.0 + pow((p->clc / p->leffCV), p->cle);
    T0 = (TRatio - 1.0);
    PowWeffWr = pow(p->weffCJ * 1.0e6, p->wr) * nf;
    T1 = T2 = T3 = T4 = 0.0;
    if (M(tempMod) == 0) {
      p->ua = p->ua + p->ua1 * T0;
      p->ub = p->ub + p->ub1 * T0;
      p->uc = p->uc + p->uc1 * T0;
      p->ud = p->ud + p->ud1 * T0;
      p->vsattemp = p->vsat - p->at * T0;
      T10 = p->prt * T0;
      if (M(rdsMod)) {
        T1 = p->rdw + T10;
        T2 = M(rdwmin) + T10;
        T3 = p->rsw + T10;
        T4 = M(rswmin) + T10;
      }
      p->rds0 = (p->rdsw + T10) * nf / PowWeffWr;
      p->rdswmin = (M(rdswmin) + T10) * nf / PowWeffWr;
    } else {
      p->ua = p->ua * (1.0 + p->ua1 * delTemp);
      p->ub = p->ub * (1.0 + p->ub1 * delTemp);
      p->uc = p->uc * (1.0 + p->uc1 * delTemp);
      p->ud = p->ud * (1.0 + p->ud1 * delTemp);
      p->vsattemp = p->vsat * (1.0 - p->at * delTemp);
      T10 = 1.0 + p->prt * delTemp;
      if (M(rdsMod)) {
        T1 = p->rdw * T10;
        T2 = M(rdwmin) * T10;
        T3 = p->rsw * T10;
        T4 = M(rswmin) * T10;
      }
      p->rds0 = p->rdsw * T10 * nf / PowWeffWr;
      p->rdswmin = M(rdswmin) * T10 * nf / PowWeffWr;
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
    if ((MG(vtl)) && (M(vtl) > 0.0)) {
      if (M(lc) < 0.0) p->lc = 0.0;
      else p->lc = M(lc);
      T0 = p->leff / (p->xn * p->leff + p->lc);
      p->tfactor = (1.0 - T0) / (1.0 + T0);
    }
    p->cgdo = (M(cgdo) + p->cf) * p->weffCV;
    p->cgso = (M(cgso) + p->cf) * p->weffCV;
    p->cgbo = M(cgbo) * p->leffCV * nf;
    if (!MG(ndep) && MG(gamma1)) {
      T0 = p->gamma1 * m->coxe;
      p->ndep = 3.01248e22 * T0 * T0;
    }
    p->phi = Vtm0 * log(p->ndep / ni) + p->phin + 0.4;
    p->sqrtPhi = sqrt(p->phi);
    p->phis3 = p->sqrtPhi * p->phi;
    p->Xdep0 = sqrt(2.0 * epssub / (B4_Charge_q * p->ndep * 1.0e6)) * p->sqrtPhi;
    p->sqrtXdep0 = sqrt(p->Xdep0);
    if (M(mtrlMod) == 0) p->litl = sqrt(3.0 * p->xj * toxe);
    else p->litl = sqrt(M(epsrsub) / epsrox * p->xj * toxe);
    p->vbi = Vtm0 * log(p->nsd * p->ndep / (ni * ni));
    if (M(mtrlMod) == 0) {
      if (p->ngate > 0.0) p->vfbsd = Vtm0 * log(p->ngate / p->nsd);
      else p->vfbsd = 0.0;
    } else {
      T0 = Vtm0 * log(p->nsd / ni);
      T1 = 0.5 * Eg0;
      if (T0 > T1) T0 = T1;
      T2 = M(easub) + T1 - m->dtype * T0;
      p->vfbsd = M(phig) - T2;
    }
    p->cdep0 = sqrt(B4_Charge_q * epssub * p->ndep * 1.0e6 / 2.0 / p->phi);
    p->ToxRatio = exp(p->ntox * log(M(toxref) / toxe)) / toxe / toxe;
    p->ToxRatioEdge = exp(p->ntox * log(M(toxref) / (toxe * p->poxedge))) / toxe / toxe
                      / p->poxedge / p->poxedge;
    p->Aechvb = (m->dtype == 1) ? 4.97232e-7 : 3.42537e-7;
    p->Bechvb = (m->dtype == 1) ? 7.45669e11 : 1.16645e12;
    p->AechvbEdgeS = p->Aechvb * p->weff * M(dlcig) * p->ToxRatioEdge;
    p->AechvbEdgeD = p->Aechvb * p->weff * M(dlcigd) * p->ToxRatioEdge;
    p->BechvbEdge = -p->Bechvb * M(toxe) * p->poxedge;
    p->Aechvb *= p->weff * p->leff * p->ToxRatio;
    p->Bechvb *= -toxe;
    p->mstar = 0.5 + atan(p->minv) / 3.14159265358979323846;
    p->mstarcv = 0.5 + atan(p->minvcv) / 3.14159265358979323846;
    p->voffcbn = p->voff + M(voffl) / p->leff;
    p->voffcbncv = p->voffcv + M(voffcvl) / p->leff;
    p->ldeb = sqrt(epssub * Vtm0 / (B4_Charge_q * p->ndep * 1.0e6)) / 3.0;
    p->acde *= pow((p->ndep / 2.0e16), -0.25);
    if (MG(k1) || MG(k2)) {
      if (!MG(k1)) p->k1 = 0.53;
      if (!MG(k2)) p->k2 = -0.0186;
    } else {
      if (!MG(vbx)) p->vbx = p->phi - 7.7348e-4 * p->ndep * p->xt * p->xt;
      if (p->vbx > 0.0) p->vbx = -p->vbx;
      if (p->vbm > 0.0) p->vbm = -p->vbm;
      if (!MG(gamma1)) p->gamma1 = 5.753e-12 * sqrt(p->ndep) / m->coxe;
      if (!MG(gamma2)) p->gamma2 = 5.753e-12 * sqrt(p->nsub) / m->coxe;
      T0 = p->gamma1 - p->gamma2;
      T1 = sqrt(p->phi - p->vbx) - p->sqrtPhi;
      T2 = sqrt(p->phi * (p->phi - p->vbm)) - p->phi;
      p->k2 = T0 * T1 / (2.0 * T2 + p->vbm);
      p->k1 = p->gamma2 - 2.0 * p->k2 * sqrt(p->phi - p->vbm);
    }
    if (!MG(vfb)) {
      if (MG(vth0)) p->vfb = m->dtype * p->vth0 - p->phi - p->k1 * p->sqrtPhi;
      else if ((M(mtrlMod)) && (MG(phig)) && (MG(nsub))) {
        T0 = Vtm0 * log(p->nsub / ni);
        T1 = 0.5 * Eg0;
        if (T0 > T1) T0 = T1;
        T2 = M(easub) + T1 + m->dtype * T0;
        p->vfb = M(phig) - T2;
      } else {
        p->vfb = -1.0;
      }
    }
    if (!MG(vth0)) p->vth0 = m->dtype * (p->vfb + p->phi + p->k1 * p->sqrtPhi);
    p->k1ox = p->k1 * toxe / M(toxm);
    tmp = sqrt(epssub / (epsrox * B4_EPS0) * toxe * p->Xdep0);
    T0 = p->dsub * p->leff / tmp;
    if (T0 < OR_EXP_THRESHOLD) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
      p->theta0vb0 = T1 / T4;
    } else {
      p->theta0vb0 = 1.0 / (OR_MAX_EXP - 2.0);
    }
    T0 = p->drout * p->leff / tmp;
    if (T0 < OR_EXP_THRESHOLD) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
      T5 = T1 / T4;
    } else {
      T5 = 1.0 / (OR_MAX_EXP - 2.0);
    }
    p->thetaRout = p->pdibl1 * T5 + p->pdibl2;
    tmp = sqrt(p->Xdep0);
    tmp1 = p->vbi - p->phi;
    tmp2 = m->factor1 * tmp;
    T0 = p->dvt1w * p->weff * p->leff / tmp2;
    if (T0 < OR_EXP_THRESHOLD) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
      T8 = T1 / T4;
    } else {
      T8 = 1.0 / (OR_MAX_EXP - 2.0);
    }
    T0 = p->dvt0w * T8;
    T8 = T0 * tmp1;
    T0 = p->dvt1 * p->leff / tmp2;
    if (T0 < OR_EXP_THRESHOLD) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
      T9 = T1 / T4;
    } else {
      T9 = 1.0 / (OR_MAX_EXP - 2.0);
    }
    T9 = p->dvt0 * T9 * tmp1;
    T4 = toxe * p->phi / (p->weff + p->w0);
    T0 = sqrt(1.0 + p->lpe0 / p->leff);
    T3 = 0.0;
    if ((M(tempMod) == 1) || (M(tempMod) == 0)) T3 = (p->kt1 + p->kt1l / p->leff) * (TRatio - 1.0);
    if (M(tempMod) == 2) T3 = -p->kt1 * (TRatio - 1.0);
    T5 = p->k1ox * (T0 - 1.0) * p->sqrtPhi + T3;
    p->vfbzbfactor = -T8 - T9 + p->k3 * T4 + T5 - p->phi - p->k1 * p->sqrtPhi;
    wlod = M(wlod);
    if (M(wlod) < 0.0) wlod = 0.0;
    T0 = pow(Lnew, M(llodku0));
    W_tmp = Wnew + wlod;
    T1 = pow(W_tmp, M(wlodku0));
    tmp1 = M(lku0) / T0 + M(wku0) / T1 + M(pku0) / (T0 * T1);
    p->ku0 = 1.0 + tmp1;
    T0 = pow(Lnew, M(llodvth));
    T1 = pow(W_tmp, M(wlodvth));
    tmp1 = M(lkvth0) / T0 + M(wkvth0) / T1 + M(pkvth0) / (T0 * T1);
    p->kvth0 = 1.0 + tmp1;
    p->kvth0 = sqrt(p->kvth0 * p->kvth0 + B4_DELTA);
    T0 = (TRatio - 1.0);
    p->ku0temp = p->ku0 * (1.0 + M(tku0) * T0) + B4_DELTA;
    Inv_saref = 1.0 / (M(saref) + 0.5 * Ldrn);
    Inv_sbref = 1.0 / (M(sbref) + 0.5 * Ldrn);
    p->inv_od_ref = Inv_saref + Inv_sbref;
    p->rho_ref = M(ku0) / p->ku0temp * p->inv_od_ref;
  }
  I->paramPtr = p;

  if ((I->sa > 0.0) && (I->sb > 0.0) && ((nf == 1.0) || ((nf > 1.0) && (I->sd > 0.0)))) {
    int i;
    Inv_sa = 0;
    Inv_sb = 0;
    kvsat = M(kvsat);
    if (M(kvsat) < -1.0) kvsat = -1.0;
    if (M(kvsat) > 1.0) kvsat = 1.0;
    for (i = 0; i < nf; i++) {
      T0 = 1.0 / nf / (I->sa + 0.5 * Ldrn + i * (I->sd + Ldrn));
      T1 = 1.0 / nf / (I->sb + 0.5 * Ldrn + i * (I->sd + Ldrn));
      Inv_sa += T0;
      Inv_sb += T1;
    }
    Inv_ODeff = Inv_sa + Inv_sb;
    rho = M(ku0) / p->ku0temp * Inv_ODeff;
    T0 = (1.0 + rho) / (1.0 + p->rho_ref);
    I->u0temp = p->u0temp * T0;
    T1 = (1.0 + kvsat * rho) / (1.0 + kvsat * p->rho_ref);
    I->vsattemp = p->vsattemp * T1;
    OD_offset = Inv_ODeff - p->inv_od_ref;
    dvth0_lod = M(kvth0) / p->kvth0 * OD_offset;
    dk2_lod = M(stk2) / pow(p->kvth0, M(lodk2)) * OD_offset;
    deta0_lod = M(steta0) / pow(p->kvth0, M(lodeta0)) * OD_offset;
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
  if (M(wpemod)) {
    if ((!I->sca_given) && (!I->scb_given) && (!I->scc_given)) {
      if ((I->sc_given) && (I->sc > 0.0)) {
        T1 = I->sc + Wdrn;
        T2 = 1.0 / M(scref);
        I->sca = M(scref) * M(scref) / (I->sc * T1);
        I->scb = ((0.1 * I->sc + 0.01 * M(scref)) * exp(-10.0 * I->sc * T2)
                  - (0.1 * T1 + 0.01 * M(scref)) * exp(-10.0 * T1 * T2)) / Wdrn;
        I->scc = ((0.05 * I->sc + 0.0025 * M(scref)) * exp(-20.0 * I->sc * T2)
                  - (0.05 * T1 + 0.0025 * M(scref)) * exp(-20.0 * T1 * T2)) / Wdrn;
      }
    }
    sceff = I->sca + M(web) * I->scb + M(wec) * I->scc;
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
  I->k2ox = I->k2 * toxe / M(toxm);
  I->vfbzb = p->vfbzbfactor + m->dtype * I->vth0;
  I->cgso = p->cgso;
  I->cgdo = p->cgdo;
  lnl = log(p->leff * 1.0e6);
  lnw = log(p->weff * 1.0e6);
  lnnf = log(nf);
  bodymode = 5;
  if ((!MG(rbps0)) || (!MG(rbpd0))) bodymode = 1;
  else if ((!MG(rbsbx0) && !MG(rbsby0)) || (!MG(rbdbx0) && !MG(rbdby0))) bodymode = 3;
  if (I->irbodyMod == 2) {
    if (bodymode == 5) {
      rbsbx = exp(log(M(rbsbx0)) + M(rbsdbxl) * lnl + M(rbsdbxw) * lnw + M(rbsdbxnf) * lnnf);
      rbsby = exp(log(M(rbsby0)) + M(rbsdbyl) * lnl + M(rbsdbyw) * lnw + M(rbsdbynf) * lnnf);
      I->rbsb = rbsbx * rbsby / (rbsbx + rbsby);
      rbdbx = exp(log(M(rbdbx0)) + M(rbsdbxl) * lnl + M(rbsdbxw) * lnw + M(rbsdbxnf) * lnnf);
      rbdby = exp(log(M(rbdby0)) + M(rbsdbyl) * lnl + M(rbsdbyw) * lnw + M(rbsdbynf) * lnnf);
      I->rbdb = rbdbx * rbdby / (rbdbx + rbdby);
    }
    if ((bodymode == 3) || (bodymode == 5)) {
      I->rbps = exp(log(M(rbps0)) + M(rbpsl) * lnl + M(rbpsw) * lnw + M(rbpsnf) * lnnf);
      I->rbpd = exp(log(M(rbpd0)) + M(rbpdl) * lnl + M(rbpdw) * lnw + M(rbpdnf) * lnnf);
    }
    rbpbx = exp(log(M(rbpbx0)) + M(rbpbxl) * lnl + M(rbpbxw) * lnw + M(rbpbxnf) * lnnf);
    rbpby = exp(log(M(rbpby0)) + M(rbpbyl) * lnl + M(rbpbyw) * lnw + M(rbpbynf) * lnnf);
    I->rbpb = rbpbx * rbpby / (rbpbx + rbpby);
  }
  if ((I->irbodyMod == 1) || ((I->irbodyMod == 2) && (bodymode == 5))) {
    I->grbdb = (I->rbdb < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbdb;
    I->grbpb = (I->rbpb < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbpb;
    I->grbps = (I->rbps < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbps;
    I->grbsb = (I->rbsb < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbsb;
    I->grbpd = (I->rbpd < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbpd;
  }
  if ((I->irbodyMod == 2) && (bodymode == 3)) {
    I->grbdb = I->grbsb = M(gbmin);
    I->grbpb = (I->rbpb < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbpb;
    I->grbps = (I->rbps < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbps;
    I->grbpd = (I->rbpd < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbpd;
  }
  if ((I->irbodyMod == 2) && (bodymode == 1)) {
    I->grbdb = I->grbsb = M(gbmin);
    I->grbps = I->grbpd = 1.0e3;
    I->grbpb = (I->rbpb < 1.0e-3) ? 1.0e3 : M(gbmin) + 1.0 / I->rbpb;
  }
  I->grgeltd = M(rshg) * (I->xgw + p->weffCJ / 3.0 / I->ngcon) / (I->ngcon * nf * (Lnew - M(xgl)));
  if (I->grgeltd > 0.0) I->grgeltd = 1.0 / I->grgeltd;
  else I->grgeltd = 1.0e3;
  I->DMCGeff = M(dmcg) - M(dmcgt);
  I->DMCIeff = M(dmci);
  I->DMDGeff = M(dmdg) - M(dmcgt);
  if (I->sourcePerimeter_given) {
    if (M(perMod) == 0) I->Pseff = I->sourcePerimeter;
    else I->Pseff = I->sourcePerimeter - p->weffCJ * nf;
  } else {
    PAeffGeo(nf, I->igeoMod, I->imin, p->weffCJ, I->DMCGeff, I->DMCIeff, I->DMDGeff,
             &I->Pseff, &dumPd, &dumAs, &dumAd);
  }
  if (I->drainPerimeter_given) {
    if (M(perMod) == 0) I->Pdeff = I->drainPerimeter;
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
  if (M(rdsMod) || I->sourceSquares_given) {
    I->sourceConductance = 0.0;
    if (I->sourceSquares_given) I->sourceConductance = M(sheetResistance) * I->sourceSquares;
    else if (I->irgeoMod > 0)
      RdseffGeo(nf, I->igeoMod, I->irgeoMod, I->imin, p->weffCJ, M(sheetResistance),
                I->DMCGeff, I->DMCIeff, I->DMDGeff, 1, &I->sourceConductance);
    else I->sourceConductance = 0.0;
    if (I->sourceConductance > 0.0) I->sourceConductance = 1.0 / I->sourceConductance;
    else I->sourceConductance = 1.0e3;
  } else {
    I->sourceConductance = 0.0;
  }
  if (M(rdsMod) || I->drainSquares_given) {
    I->drainConductance = 0.0;
    if (I->drainSquares_given) I->drainConductance = M(sheetResistance) * I->drainSquares;
    else if (I->irgeoMod > 0)
      RdseffGeo(nf, I->igeoMod, I->irgeoMod, I->imin, p->weffCJ, M(sheetResistance),
                I->DMCGeff, I->DMCIeff, I->DMDGeff, 0, &I->drainConductance);
    else I->drainConductance = 0.0;
    if (I->drainConductance > 0.0) I->drainConductance = 1.0 / I->drainConductance;
    else I->drainConductance = 1.0e3;
  } else {
    I->drainConductance = 0.0;
  }
  Nvtms = m->vtm * M(SjctEmissionCoeff);
  if ((I->Aseff <= 0.0) && (I->Pseff <= 0.0)) SourceSatCurrent = 1.0e-14;
  else SourceSatCurrent = I->Aseff * m->SjctTempSatCurDensity + I->Pseff * m->SjctSidewallTempSatCurDensity
                          + p->weffCJ * nf * m->SjctGateSidewallTempSatCurDensity;
  if (SourceSatCurrent > 0.0) {
    switch ((int)M(dioMod)) {
      case 0:
        if ((M(bvs) / Nvtms) > OR_EXP_THRESHOLD) I->XExpBVS = M(xjbvs) * OR_MIN_EXP;
        else I->XExpBVS = M(xjbvs) * exp(-M(bvs) / Nvtms);
        break;
      case 1:
        DioIjthVjmEval(Nvtms, M(ijthsfwd), SourceSatCurrent, 0.0, &I->vjsmFwd);
        I->IVjsmFwd = SourceSatCurrent * exp(I->vjsmFwd / Nvtms);
        break;
      case 2:
        if ((M(bvs) / Nvtms) > OR_EXP_THRESHOLD) {
          I->XExpBVS = M(xjbvs) * OR_MIN_EXP;
          tmp = OR_MIN_EXP;
        } else {
          I->XExpBVS = exp(-M(bvs) / Nvtms);
          tmp = I->XExpBVS;
          I->XExpBVS *= M(xjbvs);
        }
        DioIjthVjmEval(Nvtms, M(ijthsfwd), SourceSatCurrent, I->XExpBVS, &I->vjsmFwd);
        T0 = exp(I->vjsmFwd / Nvtms);
        I->IVjsmFwd = SourceSatCurrent * (T0 - I->XExpBVS / T0 + I->XExpBVS - 1.0);
        I->SslpFwd = SourceSatCurrent * (T0 + I->XExpBVS / T0) / Nvtms;
        T2 = M(ijthsrev) / SourceSatCurrent;
        if (T2 < 1.0) T2 = 10.0;
        I->vjsmRev = -M(bvs) - Nvtms * log((T2 - 1.0) / M(xjbvs));
        T1 = M(xjbvs) * exp(-(M(bvs) + I->vjsmRev) / Nvtms);
        I->IVjsmRev = SourceSatCurrent * (1.0 + T1);
        I->SslpRev = -SourceSatCurrent * T1 / Nvtms;
        break;
      default: break;
    }
  }
  Nvtmd = m->vtm * M(DjctEmissionCoeff);
  if ((I->Adeff <= 0.0) && (I->Pdeff <= 0.0)) DrainSatCurrent = 1.0e-14;
  else DrainSatCurrent = I->Adeff * m->DjctTempSatCurDensity + I->Pdeff * m->DjctSidewallTempSatCurDensity
                         + p->weffCJ * nf * m->DjctGateSidewallTempSatCurDensity;
  if (DrainSatCurrent > 0.0) {
    switch ((int)M(dioMod)) {
      case 0:
        if ((M(bvd) / Nvtmd) > OR_EXP_THRESHOLD) I->XExpBVD = M(xjbvd) * OR_MIN_EXP;
        else I->XExpBVD = M(xjbvd) * exp(-M(bvd) / Nvtmd);
        break;
      case 1:
        DioIjthVjmEval(Nvtmd, M(ijthdfwd), DrainSatCurrent, 0.0, &I->vjdmFwd);
        I->IVjdmFwd = DrainSatCurrent * exp(I->vjdmFwd / Nvtmd);
        break;
      case 2:
        if ((M(bvd) / Nvtmd) > OR_EXP_THRESHOLD) {
          I->XExpBVD = M(xjbvd) * OR_MIN_EXP;
          tmp = OR_MIN_EXP;
        } else {
          I->XExpBVD = exp(-M(bvd) / Nvtmd);
          tmp = I->XExpBVD;
          I->XExpBVD *= M(xjbvd);
        }
        DioIjthVjmEval(Nvtmd, M(ijthdfwd), DrainSatCurrent, I->XExpBVD, &I->vjdmFwd);
        T0 = exp(I->vjdmFwd / Nvtmd);
        I->IVjdmFwd = DrainSatCurrent * (T0 - I->XExpBVD / T0 + I->XExpBVD - 1.0);
        I->DslpFwd = DrainSatCurrent * (T0 + I->XExpBVD / T0) / Nvtmd;
        T2 = M(ijthdrev) / DrainSatCurrent;
        if (T2 < 1.0) T2 = 10.0;
        I->vjdmRev = -M(bvd) - Nvtmd * log((T2 - 1.0) / M(xjbvd));
        T1 = M(xjbvd) * exp(-(M(bvd) + I->vjdmRev) / Nvtmd);
        I->IVjdmRev = DrainSatCurrent * (1.0 + T1);
        I->DslpRev = -DrainSatCurrent * T1 / Nvtmd;
        break;
      default: break;
    }
  }
  T0 = (TRatio - 1.0);
  m->njtsstemp = M(njts) * (1.0 + M(tnjts) * T0);
  m->njtsswstemp = M(njtssw) * (1.0 + M(tnjtssw) * T0);
  m->njtsswgstemp = M(njtsswg) * (1.0 + M(tnjtsswg) * T0);
  m->njtsdtemp = M(njtsd) * (1.0 + M(tnjtsd) * T0);
  m->njtsswdtemp = M(njtsswd) * (1.0 + M(tnjtsswd) * T0);
  m->njtsswgdtemp = M(njtsswgd) * (1.0 + M(tnjtsswgd) * T0);
  T7 = Eg0 / m->vtm * T0;
#define OR_DEXP2(A, B) do { if ((A) > OR_EXP_THRESHOLD) (B) = OR_MAX_EXP * (1.0 + (A) - OR_EXP_THRESHOLD); \
    else if ((A) < -OR_EXP_THRESHOLD) (B) = OR_MIN_EXP; else (B) = exp(A); } while (0)
  T9 = M(xtss) * T7;
  OR_DEXP2(T9, T1);
  T9 = M(xtsd) * T7;
  OR_DEXP2(T9, T2);
  T9 = M(xtssws) * T7;
  OR_DEXP2(T9, T3);
  T9 = M(xtsswd) * T7;
  OR_DEXP2(T9, T4);
  T9 = M(xtsswgs) * T7;
  OR_DEXP2(T9, T5);
  T9 = M(xtsswgd) * T7;
  OR_DEXP2(T9, T6);
  T10 = p->weffCJ * nf;
  I->SjctTempRevSatCur = T1 * I->Aseff * M(jtss);
  I->DjctTempRevSatCur = T2 * I->Adeff * M(jtsd);
  I->SswTempRevSatCur = T3 * I->Pseff * M(jtssws);
  I->DswTempRevSatCur = T4 * I->Pdeff * M(jtsswd);
  I->SswgTempRevSatCur = T5 * T10 * M(jtsswgs);
  I->DswgTempRevSatCur = T6 * T10 * M(jtsswgd);
  if (M(mtrlMod)) {
    /* TOXP from EOT (:7614-7716): Vgs_eff at VDDEOT with poly depletion,
     * Vth at Vds = Vbs = 0, then the charge centroid iterated */
    double vddeot, Vgs_eff, V0, lt1, ltw, Theta0, Delt_vth, TempRatio, Vth_NarrowW, Lpe_Vb, Vth;
    double n, tmp3, Vgsteff, toxpf, toxpi, Tcen;
    int niter;
    tmp2 = I->vfb + p->phi;
    vddeot = m->dtype * M(vddeot);
    T0 = M(epsrgate) * B4_EPS0;
    if ((p->ngate > 1.0e18) && (p->ngate < 1.0e25) && (vddeot > tmp2) && (T0 != 0)) {
      T1 = 1.0e6 * OR_Q * T0 * p->ngate / (m->coxe * m->coxe);
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
    if (T0 < OR_EXP_THRESHOLD) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
      Theta0 = T1 / T4;
    } else {
      Theta0 = 1.0 / (OR_MAX_EXP - 2.0);
    }
    Delt_vth = p->dvt0 * Theta0 * V0;
    T0 = p->dvt1w * p->weff * p->leff / ltw;
    if (T0 < OR_EXP_THRESHOLD) {
      T1 = exp(T0);
      T2 = T1 - 1.0;
      T3 = T2 * T2;
      T4 = T3 + 2.0 * T1 * OR_MIN_EXP;
      T5 = T1 / T4;
    } else {
      T5 = 1.0 / (OR_MAX_EXP - 2.0);
    }
    T2 = p->dvt0w * T5 * V0;
    TempRatio = I->temp / M(tnom) - 1.0;
    T0 = sqrt(1.0 + p->lpe0 / p->leff);
    T1 = p->k1ox * (T0 - 1.0) * p->sqrtPhi + (p->kt1 + p->kt1l / p->leff) * TempRatio;
    Vth_NarrowW = toxe * p->phi / (p->weff + p->w0);
    Lpe_Vb = sqrt(1.0 + p->lpeb / p->leff);
    Vth = m->dtype * I->vth0 + (p->k1ox - p->k1) * p->sqrtPhi * Lpe_Vb - Delt_vth - T2
          + p->k3 * Vth_NarrowW + T1;
    tmp1 = epssub / p->Xdep0;
    I->nstar = m->vtm / B4_Charge_q * (m->coxe + tmp1 + p->cit);
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
      if (M(tempMod) < 2) T4 = m->vtm * log(p->leff / T3);
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
      T1 = 1.0 + exp(M(bdos) * 0.7 * log(T0));
      Tcen = M(ados) * 1.9e-9 / T1;
      toxpf = toxe - epsrox / M(epsrsub) * Tcen;
      niter++;
    } while ((niter <= 4) && (fabs(toxpf - toxpi) > 1e-12));
    M(toxp) = toxpf;
    m->coxp = epsrox * B4_EPS0 / M(toxp);
  }
}

/* Instance ctor defaults (N_DEV_MOSFET_B4.C:4585-4618) and
 * Instance::processParams (:3621-3794).  Slots of g->ival follow the B4I
 * order of the parameter table. */
int b4_instance_init(struct or_group_s* g, b4inst* I, int64_t i) {
  b4model* m = I->model;
  int k = 0;
  double Rtot;
  int createNode;
#define B4I(NAME, mem, def) I->mem = g->ival[k][i]; I->mem##_given = g->igiven[k]; ++k;
#include "../include/xa_bsim4_params.def"
#undef B4I
  if (!I->rbdb_given) I->rbdb = M(rbdb);
  if (!I->rbsb_given) I->rbsb = M(rbsb);
  if (!I->rbpb_given) I->rbpb = M(rbpb);
  if (!I->rbps_given) I->rbps = M(rbps);
  if (!I->rbpd_given) I->rbpd = M(rbpd);
  if (!I->xgw_given) I->xgw = M(xgw);
  if (!I->ngcon_given) I->ngcon = M(ngcon);
  if (!I->sd_given) I->sd = 2.0 * M(dmcg);
  if (!I->temp_given) I->temp = g->opt.temp;
  if (!I->drainArea_given) I->drainArea = g->opt.defad;
  if (!I->sourceArea_given) I->sourceArea = g->opt.defas;
  I->irbodyMod = (int)I->rbodyMod;
  I->irgateMod = (int)I->rgateMod;
  I->igeoMod = (int)I->geoMod;
  I->irgeoMod = (int)I->rgeoMod;
  I->itrnqsMod = (int)I->trnqsMod;
  I->iacnqsMod = (int)I->acnqsMod;
  I->imin = (int)I->min;
  if (!I->rbodyMod_given) I->irbodyMod = (int)M(rbodyMod);
  else if (I->irbodyMod != 0 && I->irbodyMod != 1 && I->irbodyMod != 2) I->irbodyMod = (int)M(rbodyMod);
  if (!I->rgateMod_given) I->irgateMod = (int)M(rgateMod);
  else if (I->irgateMod < 0 || I->irgateMod > 3) I->irgateMod = (int)M(rgateMod);
  if (!I->geoMod_given) I->igeoMod = (int)M(geoMod);
  if (!I->rgeoMod_given) I->irgeoMod = 0;
  if (!I->trnqsMod_given) I->itrnqsMod = (int)M(trnqsMod);
  else if (I->itrnqsMod != 0 && I->itrnqsMod != 1) I->itrnqsMod = (int)M(trnqsMod);
  if (!I->acnqsMod_given) I->iacnqsMod = (int)M(acnqsMod);
  else if (I->iacnqsMod != 0 && I->iacnqsMod != 1) I->iacnqsMod = (int)M(acnqsMod);
  b4_update_temperature(I, I->temp, &g->opt);
  createNode = 0;
  if (M(rdsMod) != 0) createNode = 1;
  else if (M(sheetResistance) > 0) {
    if (I->drainSquares_given && I->drainSquares > 0) createNode = 1;
    else if (!I->drainSquares_given && (I->irgeoMod != 0)) {
      RdseffGeo(I->nf, I->igeoMod, I->irgeoMod, I->imin, I->w, M(sheetResistance), I->DMCGeff,
                I->DMCIeff, I->DMDGeff, 0, &Rtot);
      if (Rtot > 0) createNode = 1;
    }
  }
  I->drainMOSFET_B4Exists = createNode != 0;
  createNode = 0;
  if (M(rdsMod) != 0) createNode = 1;
  else if (M(sheetResistance) > 0) {
    if (I->sourceSquares_given && I->sourceSquares > 0) createNode = 1;
    else if (!I->sourceSquares_given && (I->irgeoMod != 0)) {
      RdseffGeo(I->nf, I->igeoMod, I->irgeoMod, I->imin, I->w, M(sheetResistance), I->DMCGeff,
                I->DMCIeff, I->DMDGeff, 1, &Rtot);
      if (Rtot > 0) createNode = 1;
    }
  }
  I->sourceMOSFET_B4Exists = createNode != 0;
  /* numIntVars / numStateVars (:3754-3788) */
  g->n_ext[i] = 4;
  g->n_int[i] = (I->drainMOSFET_B4Exists ? 1 : 0) + (I->sourceMOSFET_B4Exists ? 1 : 0)
                + ((I->irgateMod == 1 || I->irgateMod == 2) ? 1 : (I->irgateMod == 3) ? 2 : 0)
                + (I->itrnqsMod ? 1 : 0) + (I->irbodyMod ? 3 : 0)
                + (I->icVBS_given ? 1 : 0) + (I->icVDS_given ? 1 : 0) + (I->icVGS_given ? 1 : 0);
  g->n_sta[i] = 3 + (I->irgateMod == 3 ? 1 : 0) + (I->irbodyMod ? 2 : 0) + (I->itrnqsMod ? 2 : 0);
  g->n_sto[i] = 14;   /* the BSIM4 Master loads write no lead currents */
  I->ScalingFactor = 1.0e-9;
  return XA_OK;
}
