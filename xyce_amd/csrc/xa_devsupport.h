// xa_devsupport.h -- SPICE voltage limiting and Meyer capacitance helpers,
// evaluated per instance inside the device kernels.
//
// Behaviour follows DeviceSupport (src/DeviceModelPKG/Core/src/
// N_DEV_DeviceSupport.C): lambertw :105-176, limvds :186-203, pnjlim :214-252,
// fetlim :337-420, qmeyer :545-618, qmeyerderivs :633-742.  Constants follow
// src/DeviceModelPKG/Core/include/N_DEV_Const.h.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

// Parity build (XA_PORTABLE_MATH): exp, log, pow, tan and sin in the evaluation
// code from here on -- these helpers and every model file that includes
// this header -- are the portable functions of xa_pmath.h, which the CPU
// restatement's parity build uses too, so the two agree bit for bit.
#ifdef XA_PORTABLE_MATH
#include "xa_pmath.h"
#define exp(a) xa_pm_exp(a)
#define log(a) xa_pm_log(a)
#define pow(a, b) xa_pm_pow(a, b)
#define tan(a) xa_pm_tan(a)
#define sin(a) xa_pm_sin(a)
#endif

#define XA_HD __host__ __device__ __forceinline__

namespace xa {

// Physical constants (N_DEV_Const.h:50-114).
constexpr double kQ        = 1.6021918e-19;
constexpr double kCtoK     = 273.15;
constexpr double kBoltz    = 1.3806226e-23;
constexpr double kRefTemp  = 300.15;
constexpr double kKoverQ   = kBoltz / kQ;
constexpr double kEg300    = 1.1150877;
constexpr double kEg0      = 1.16;
constexpr double kAlphaEg  = 7.02e-4;
constexpr double kBetaEg   = 1108.0;
constexpr double kNi0      = 1.45e10;
constexpr double kEpsOx    = 3.453133e-11;
constexpr double kEpsSi    = 1.03594e-10;
constexpr double kMaxExp   = 5.834617425e14;
constexpr double kMinExp   = 1.713908431e-15;
constexpr double kExpThreshold = 34.0;
constexpr double kMaxExpArg = 100.0;
// sqrt(2) as a correctly rounded literal (the reference's sqrt(2.0); sqrt is
// correctly rounded on both sides).
constexpr double kRoot2    = 1.4142135623730951;
// N_DEV_Const.h:60 defines CONSTe as exp(1.0), a math-library call; the loads
// call it the same way, so e is whatever this build's exp gives.
#define XA_CONSTE exp(1.0)

XA_HD double dmin(double a, double b) { return a < b ? a : b; }
XA_HD double dmax(double a, double b) { return a > b ? a : b; }

// Per-iteration limit on V_DS (N_DEV_DeviceSupport.C:186-203).
XA_HD double limvds(double vnew, double vold) {
  if (vold >= 3.5) {
    if (vnew > vold) return dmin(vnew, 3.0 * vold + 2.0);
    if (vnew < 3.5) return dmax(vnew, 2.0);
    return vnew;
  }
  return (vnew > vold) ? dmin(vnew, 4.0) : dmax(vnew, -0.5);
}

// Log-damped PN-junction limiting (N_DEV_DeviceSupport.C:214-252).
XA_HD double pnjlim(double vnew, double vold, double vt, double vcrit, int* limited) {
  if (vnew > vcrit && fabs(vnew - vold) > vt + vt) {
    if (vold > 0) {
      const double arg = 1 + (vnew - vold) / vt;
      vnew = (arg > 0) ? vold + vt * log(arg) : vcrit;
    } else {
      vnew = vt * log(vnew / vt);
    }
    *limited = 1;
  } else {
    *limited = 0;
  }
  return vnew;
}

// FET gate-voltage limiting around the threshold (N_DEV_DeviceSupport.C:337-420).
XA_HD double fetlim(double vnew, double vold, double vto) {
  const double hi   = fabs(2 * (vold - vto)) + 2;
  const double lo   = hi / 2 + 2;
  const double vtox = vto + 3.5;
  const double dv   = vnew - vold;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (dv <= 0) {                       // turning off
        if (vnew >= vtox) { if (-dv > lo) vnew = vold - lo; }
        else              { vnew = dmax(vnew, vto + 2.0); }
      } else if (dv >= hi) {               // staying on
        vnew = vold + hi;
      }
    } else {                               // middle region
      vnew = (dv <= 0) ? dmax(vnew, vto - 0.5) : dmin(vnew, vto + 4.0);
    }
  } else {                                 // off
    if (dv <= 0) {
      if (-dv > hi) vnew = vold - hi;
    } else {
      const double vtemp = vto + 0.5;
      if (vnew <= vtemp) { if (dv > lo) vnew = vold + lo; }
      else               { vnew = vtemp; }
    }
  }
  return vnew;
}

// Lambert W by Halley-type iteration from a regional initial guess
// (Barry et al. / Day; N_DEV_DeviceSupport.C:105-176).
XA_HD void lambertw(double x, double& w, int& ierr, double& xi) {
  const double turnpt = -0.36787944117144233;   // -exp(-1)
  const double c1 = 1.5, c2 = 0.75, eps = 2.e-15;
  double relerr = 1.0;
  int it = 0;
  ierr = 0;
  if (x > c1) {
    w = c2 * log(x);
    xi = log(x / w) - w;
  } else if (x >= 0.0) {
    w = x;
    if (x == 0.0) return;
    if (x < 1 - c2) w = x * (1. - x + c1 * x * x);
    xi = -w;
  } else if (x >= turnpt) {
    if (x > -0.2) {
      w = x * (1.0 - x + c1 * x * x);
      xi = log(1.0 - x + c1 * x * x) - w;
    } else {
      const double d = fabs(x - turnpt);
      w = -1 + sqrt(2.0 * XA_CONSTE) * sqrt(x - turnpt);
      if (d == 0.0) return;
      xi = log(x / w) - w;
    }
  } else {
    ierr = 1;
    w = -1.0;
    return;
  }
  while (relerr > eps && it < 10) {
    const double wp1 = w + 1.0;
    const double r = xi / wp1, r2 = r * r, r3 = r2 * r;
    const double s = 6. * wp1 * wp1;
    w = w * (1.0 + r + r2 / (2.0 * wp1) - (2. * w - 1.0) * r3 / s);
    if (w * x < 0.0) w = -w;
    xi = log(x / w) - w;
    relerr = fabs(x > 1.0 ? xi / w : xi);
    ++it;
  }
  if (it == 10) ierr = 2;
}

// Meyer intrinsic gate capacitances, half-values (N_DEV_DeviceSupport.C:545-618).
XA_HD void qmeyer(double vgs, double vgd, double von, double vdsat,
                  double& capgs, double& capgd, double& capgb,
                  double phi, double cox) {
  const double vgst = vgs - von;
  if (vgst <= -phi) {
    capgb = cox / 2; capgs = 0; capgd = 0;
  } else if (vgst <= -phi / 2) {
    capgb = -vgst * cox / (2 * phi); capgs = 0; capgd = 0;
  } else if (vgst <= 0) {
    capgb = -vgst * cox / (2 * phi);
    capgs = vgst * cox / (1.5 * phi) + cox / 3;
    capgd = 0;
  } else {
    const double vds = vgs - vgd;
    if (vdsat <= vds) {
      capgs = cox / 3; capgd = 0; capgb = 0;
    } else {
      const double vddif  = 2.0 * vdsat - vds;
      const double vddif1 = vdsat - vds;
      const double vddif2 = vddif * vddif;
      capgd = cox * (1.0 - vdsat * vdsat / vddif2) / 3;
      capgs = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3;
      capgb = 0;
    }
  }
}

// Partial derivatives of the Meyer capacitances (N_DEV_DeviceSupport.C:633-742).
// d[0..8] = dcapgs/d{vgs,vgb,vgd}, dcapgd/d{vgs,vgb,vgd}, dcapgb/d{vgs,vgb,vgd}.
XA_HD void qmeyerderivs(double vgs, double vgd, double von, double vdsat,
                        double* d, double phi, double cox) {
  for (int k = 0; k < 9; ++k) d[k] = 0;
  const double vgst = vgs - von;
  if (vgst <= -phi) return;
  if (vgst <= -phi / 2) { d[6] = -1.0 * cox / (2.0 * phi); return; }
  if (vgst <= 0) {
    d[0] = cox / (1.5 * phi);
    d[6] = -1.0 * cox / (2.0 * phi);
    return;
  }
  const double vds = vgs - vgd;
  if (vdsat <= vds) return;
  const double vgdt = vgd - von;
  const double den = vgst + vgdt;
  const double den3 = den * den * den;
  d[0] = 4.0 / 3.0 * cox * vgdt * vgdt / den3;
  d[2] = -4.0 / 3.0 * cox * vgst * vgdt / den3;
  d[3] = -4.0 / 3.0 * cox * vgst * vgdt / den3;
  d[5] = 4.0 / 3.0 * cox * vgst * vgst / den3;
}

}  // namespace xa
