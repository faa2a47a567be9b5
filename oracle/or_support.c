/* or_support.c -- oracle restatement of DeviceSupport (test infrastructure
 * only).  src/DeviceModelPKG/Core/src/N_DEV_DeviceSupport.C:
 * lambertw :105-176, limvds :186-203, pnjlim :214-252, fetlim :337-420,
 * qmeyer :545-618, qmeyerderivs :633-742. */
#include "or_internal.h"
#include "or_eval_libm.h"

/* N_DEV_DeviceSupport.C:105-176 */
void or_lambertw(double x, double* w, int* ierr, double* xi) {
  int i = 0, maxit = 10;
  const double turnpt = -exp(-1.), c1 = 1.5, c2 = .75;
  double r, r2, r3, s, mach_eps, relerr = 1., diff;
  mach_eps = 2.e-15;
  *ierr = 0;
  if (x > c1) {
    *w = c2 * log(x);
    *xi = log(x / *w) - *w;
  } else {
    if (x >= 0.0) {
      *w = x;
      if (x == 0.) return;
      if (x < (1 - c2)) *w = x * (1. - x + c1 * x * x);
      *xi = -*w;
    } else {
      if (x >= turnpt) {
        if (x > -0.2) {
          *w = x * (1.0 - x + c1 * x * x);
          *xi = log(1.0 - x + c1 * x * x) - *w;
        } else {
          diff = x - turnpt;
          if (diff < 0.0) diff = -diff;
          *w = -1 + sqrt(2.0 * exp(1.)) * sqrt(x - turnpt);
          if (diff == 0.0) return;
          *xi = log(x / *w) - *w;
        }
      } else {
        *ierr = 1;
        *w = -1.0;
        return;
      }
    }
  }
  while (relerr > mach_eps && i < maxit) {
    r = *xi / (*w + 1.0);
    r2 = r * r;
    r3 = r2 * r;
    s = 6. * (*w + 1.0) * (*w + 1.0);
    *w = *w * (1.0 + r + r2 / (2.0 * (*w + 1.0)) - (2. * *w - 1.0) * r3 / s);
    if (*w * x < 0.0) *w = -*w;
    *xi = log(x / *w) - *w;
    if (x > 1.0) relerr = *xi / *w;
    else relerr = *xi;
    if (relerr < 0.0) relerr = -relerr;
    ++i;
  }
  if (i == maxit) *ierr = 2;
}

/* N_DEV_DeviceSupport.C:186-203 */
double or_limvds(double vnew, double vold) {
  if (vold >= 3.5) {
    if (vnew > vold) vnew = or_min(vnew, (3.0 * vold) + 2.0);
    else if (vnew < 3.5) vnew = or_max(vnew, 2.0);
  } else {
    if (vnew > vold) vnew = or_min(vnew, 4.0);
    else vnew = or_max(vnew, -0.5);
  }
  return vnew;
}

/* N_DEV_DeviceSupport.C:214-252 */
double or_pnjlim(double vnew, double vold, double vt, double vcrit, int* icheck) {
  double arg;
  if ((vnew > vcrit) && (fabs(vnew - vold) > (vt + vt))) {
    if (vold > 0) {
      arg = 1 + (vnew - vold) / vt;
      if (arg > 0) vnew = vold + vt * log(arg);
      else vnew = vcrit;
    } else {
      vnew = vt * log(vnew / vt);
    }
    *icheck = 1;
  } else {
    *icheck = 0;
  }
  return vnew;
}

/* N_DEV_DeviceSupport.C:337-420 */
double or_fetlim(double vnew, double vold, double vto) {
  double vtsthi = fabs(2 * (vold - vto)) + 2;
  double vtstlo = vtsthi / 2 + 2;
  double vtox = vto + 3.5;
  double delv = vnew - vold;
  double vtemp;
  if (vold >= vto) {
    if (vold >= vtox) {
      if (delv <= 0) {
        if (vnew >= vtox) {
          if (-delv > vtstlo) vnew = vold - vtstlo;
        } else {
          vnew = or_max(vnew, vto + 2.0);
        }
      } else {
        if (delv >= vtsthi) vnew = vold + vtsthi;
      }
    } else {
      if (delv <= 0) vnew = or_max(vnew, vto - 0.5);
      else vnew = or_min(vnew, vto + 4.0);
    }
  } else {
    if (delv <= 0) {
      if (-delv > vtsthi) vnew = vold - vtsthi;
    } else {
      vtemp = vto + 0.5;
      if (vnew <= vtemp) {
        if (delv > vtstlo) vnew = vold + vtstlo;
      } else {
        vnew = vtemp;
      }
    }
  }
  return vnew;
}

/* N_DEV_DeviceSupport.C:545-618 */
void or_qmeyer(double vgs, double vgd, double vgb, double von, double vdsat,
               double* capgs, double* capgd, double* capgb, double phi, double cox) {
  double vds, vddif, vddif1, vddif2, vgst;
  (void)vgb;
  vgst = vgs - von;
  if (vgst <= -phi) {
    *capgb = cox / 2; *capgs = 0; *capgd = 0;
  } else if (vgst <= -phi / 2) {
    *capgb = -vgst * cox / (2 * phi); *capgs = 0; *capgd = 0;
  } else if (vgst <= 0) {
    *capgb = -vgst * cox / (2 * phi);
    *capgs = vgst * cox / (1.5 * phi) + cox / 3;
    *capgd = 0;
  } else {
    vds = vgs - vgd;
    if (vdsat <= vds) {
      *capgs = cox / 3; *capgd = 0; *capgb = 0;
    } else {
      vddif = 2.0 * vdsat - vds;
      vddif1 = vdsat - vds;
      vddif2 = vddif * vddif;
      *capgd = cox * (1.0 - vdsat * vdsat / vddif2) / 3;
      *capgs = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3;
      *capgb = 0;
    }
  }
}

/* N_DEV_DeviceSupport.C:633-742; d[] = gs/{gs,gb,gd}, gd/{..}, gb/{..} */
void or_qmeyerderivs(double vgs, double vgd, double vgb, double von, double vdsat,
                     double* d, double phi, double cox, int Dtype) {
  double vgst, vds, vdenom, vdenom3, vgdt;
  int k;
  (void)vgb; (void)Dtype;
  for (k = 0; k < 9; ++k) d[k] = 0;
  vgst = vgs - von;
  if (vgst <= -phi) {
  } else if (vgst <= -phi / 2) {
    d[6] = -1.0 * cox / (2.0 * phi);
  } else if (vgst <= 0) {
    d[0] = cox / (1.5 * phi);
    d[6] = -1.0 * cox / (2.0 * phi);
  } else {
    vds = vgs - vgd;
    if (vdsat <= vds) {
    } else {
      vgdt = vgd - von;
      vdenom = vgst + vgdt;
      vdenom3 = vdenom * vdenom * vdenom;
      d[0] = 4.0 / 3.0 * cox * vgdt * vgdt / vdenom3;
      d[2] = -4.0 / 3.0 * cox * vgst * vgdt / vdenom3;
      d[3] = -4.0 / 3.0 * cox * vgst * vgdt / vdenom3;
      d[5] = 4.0 / 3.0 * cox * vgst * vgst / vdenom3;
    }
  }
}
