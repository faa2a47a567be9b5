// bsim4_host.h -- BSIM4 (level 54) host-side model, size-dependent and
// instance records used while a group is set up.  The evaluation kernel
// never sees these structs: setup packs what it reads into per-instance
// SoA columns and a per-model device card.
//
// Reference: src/DeviceModelPKG/OpenModels/include/N_DEV_MOSFET_B4.h
//   SizeDependParam :88-321, Instance :322-1149, Model :1151-2138.
// Field names follow the reference's members so that each statement of the
// setup code can be read against the line it implements.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <tuple>
#include <vector>
#include "xa_devsupport.h"
#include "xa_internal.h"

namespace xa {
namespace bsim4 {

constexpr double kEps0 = 8.854214871e-12;   // N_DEV_Const.h CONSTperm0
constexpr double kEpsSi = 1.03594e-10;      // N_DEV_MOSFET_B4.C:101
constexpr double kKboQ = 8.617087e-5;       // N_DEV_MOSFET_B4.C:106
constexpr double kVt0 = kBoltz * (27.0 + kCtoK) / kQ;
constexpr double kMM = 3;
constexpr double kDelta = 1.0e-9;
constexpr double kChargeQ = 1.60219e-19;
constexpr double kDelta1 = 0.02;
constexpr double kDelta3 = 0.02;
constexpr double kDelta4 = 0.02;
constexpr double kScalingFactor = 1.0e-9;   // N_DEV_MOSFET_B4.C:7878 (NQS charge scale)

// ---- model card: every table parameter as a named field ------------------
struct Params {
#define B4M(NAME, mem, def) double mem;
#include "xa_bsim4_params.def"
#undef B4M
};
struct Given {
#define B4M(NAME, mem, def) char mem;
#include "xa_bsim4_params.def"
#undef B4M
};
// Slots of the parameter table: model rows, the TYPE pseudo-parameter (the
// nmos/pmos keyword of the .model line), then the instance rows.
enum MSlot {
#define B4M(NAME, mem, def) MS_##mem,
#include "xa_bsim4_params.def"
#undef B4M
  MS_TYPE, NM
};
enum ISlot {
#define B4I(NAME, mem, def) IS_##mem,
#include "xa_bsim4_params.def"
#undef B4I
  NI
};

// ---- size-dependent parameters (SizeDependParam) ------------------------
struct Size {
  double Length, Width, NFinger, referenceTemperature;
#define B4BIN(sm, base, l, w, p) double sm;
#include "xa_bsim4_binning.def"
#undef B4BIN
  double dl, dlc, dw, dwc, dwj, leff, weff, leffCV, weffCV, weffCJ;
  double abulkCVfactor, vsattemp, rds0, rdswmin, rd0, rdwmin, rs0, rswmin;
  double u0temp, tfactor, lc, cgdo, cgso, cgbo, phi, sqrtPhi, phis3, Xdep0, sqrtXdep0;
  double litl, vbi, vfbsd, cdep0, ToxRatio, ToxRatioEdge, Aechvb, Bechvb;
  double AechvbEdgeS, AechvbEdgeD, BechvbEdge, mstar, mstarcv, voffcbn, voffcbncv;
  double ldeb, k1ox, theta0vb0, thetaRout, vfbzbfactor, ku0, kvth0, ku0temp;
  double inv_od_ref, rho_ref;
};

// ---- model --------------------------------------------------------------
struct Model {
  Params p;
  Given g;
  int dtype = 1;   // +1 NMOS, -1 PMOS
  // Temperature-dependent values held by the model.  Every instance's
  // updateTemperature rewrites them, so the evaluation reads the values of
  // the last instance set up (as the reference's shared model object does).
  double coxe, coxp, vcrit, factor1, vtm0, vtm, Eg0;
  double SjctTempSatCurDensity, SjctSidewallTempSatCurDensity, SjctGateSidewallTempSatCurDensity;
  double DjctTempSatCurDensity, DjctSidewallTempSatCurDensity, DjctGateSidewallTempSatCurDensity;
  double SunitAreaTempJctCap, DunitAreaTempJctCap;
  double SunitLengthSidewallTempJctCap, DunitLengthSidewallTempJctCap;
  double SunitLengthGateSidewallTempJctCap, DunitLengthGateSidewallTempJctCap;
  double PhiBS, PhiBD, PhiBSWS, PhiBSWD, PhiBSWGS, PhiBSWGD;
  double njtsstemp, njtsswstemp, njtsswgstemp, njtsdtemp, njtsswdtemp, njtsswgdtemp;
  // Size-dependent records keyed on (L, W, NF): the first instance of a key
  // creates the record, later ones share it (N_DEV_MOSFET_B4.C:6018-6049).
  std::map<std::tuple<double, double, double>, std::unique_ptr<Size>> sizes;
};

// ---- instance (setup-time values) ---------------------------------------
struct Inst {
  Model* model = nullptr;
  Size* paramPtr = nullptr;
#define B4I(NAME, mem, def) double mem;
#include "xa_bsim4_params.def"
#undef B4I
#define B4I(NAME, mem, def) char mem##_given;
#include "xa_bsim4_params.def"
#undef B4I
  int irbodyMod, irgateMod, igeoMod, irgeoMod, itrnqsMod, iacnqsMod, imin;
  int drainMOSFET_B4Exists, sourceMOSFET_B4Exists;
  double u0temp, vsattemp, vth0, vfb, eta0, k2, k2ox, vtfbphi1, vtfbphi2, vbsc, vfbzb;
  double cgso, cgdo, grbdb, grbpb, grbps, grbsb, grbpd, grgeltd;
  double DMCGeff, DMCIeff, DMDGeff, Pseff, Pdeff, Aseff, Adeff;
  double sourceConductance, drainConductance;
  double XExpBVS, XExpBVD, vjsmFwd, vjsmRev, vjdmFwd, vjdmRev;
  double IVjsmFwd, IVjsmRev, IVjdmFwd, IVjdmRev, SslpFwd, SslpRev, DslpFwd, DslpRev;
  double SjctTempRevSatCur, DjctTempRevSatCur, SswTempRevSatCur, DswTempRevSatCur;
  double SswgTempRevSatCur, DswgTempRevSatCur;
};

int model_init(Model& m, const Card& c, const xa_device_options& opt);
int instance_init(Group& g, Inst& I, int64_t i);

}  // namespace bsim4
}  // namespace xa
