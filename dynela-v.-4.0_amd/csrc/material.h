// material.h -- hardening laws and prescribed-speed ramps, device side.
#pragma once

#include "engine.h"

namespace dnlg {

// JohnsonCookLaw::getYieldStress (JohnsonCookLaw.C:146-170), ElastoplasticLaw
// (ElastoplasticLaw.C:110-114) and the Voce + linear law.
__device__ __forceinline__ double law_yield(int law, const double *p, double epsp, double depsp, double T) {
  if (law == DNLG_LAW_JOHNSON_COOK) {
    const double hard = p[0] + p[1] * pow(epsp, p[3]);
    const double visc = (depsp > p[5]) ? 1.0 + p[2] * log(depsp / p[5]) : 1.0;
    double temp = 1.0;
    if (T > p[7]) temp = (T < p[6]) ? 1.0 - pow(((T - p[7]) / (p[6] - p[7])), p[4]) : 0.0;
    return hard * visc * temp;
  }
  if (law == DNLG_LAW_POWER) return p[0] + p[1] * pow(epsp, p[2]);
  return p[0] + p[1] * (1.0 - exp(-p[2] * epsp)) + p[3] * epsp;
}

// Yield stress and its consistent derivative in one pass: shares the power
// and logarithm terms of JohnsonCookLaw::getYieldStress and
// getDerivateYieldStress (JohnsonCookLaw.C:298-339).  eta_rcp = eta/(rho*Cp).
__device__ __forceinline__ void law_yield_and_slope(int law, const double *p, double eta_rcp, double epsp,
                                                    double depsp, double T, double dtime, double &yield,
                                                    double &hard) {
  if (law == DNLG_LAW_JOHNSON_COOK) {
    const double A = p[0], B = p[1], C = p[2], n = p[3], m = p[4], d0 = p[5], Tm = p[6], T0 = p[7];
    const double lg = log(epsp);
    const double pn = exp(n * lg);  // epsp^n
    const double hard0 = A + B * pn;
    const bool rate = depsp > d0;
    const double visc0 = rate ? 1.0 + C * log(depsp / d0) : 1.0;
    double tp1 = 0.0, temp0 = 1.0;
    const bool warm = T > T0;
    if (warm) {
      if (T < Tm) {
        tp1 = pow(((T - T0) / (Tm - T0)), m);
        temp0 = 1.0 - tp1;
      } else {
        temp0 = 0.0;
      }
    }
    yield = hard0 * visc0 * temp0;
    double h = n * B * (pn / epsp) * visc0 * temp0;
    if (rate) h += hard0 * C * temp0 / (depsp * dtime);
    if (warm && T < Tm) h -= m * eta_rcp * hard0 * hard0 * visc0 * visc0 * temp0 * tp1 / (T - T0);
    hard = h;
    return;
  }
  if (law == DNLG_LAW_POWER) {
    const double pn = pow(epsp, p[2]);
    yield = p[0] + p[1] * pn;
    hard = p[2] * p[1] * pow(epsp, p[2] - 1.0);
    return;
  }
  const double ex = exp(-p[2] * epsp);
  yield = p[0] + p[1] * (1.0 - ex) + p[3] * epsp;
  hard = p[1] * p[2] * ex + p[3];
}

// Per-IP constants of the radial-return iteration of the mixed path.
struct ReturnConsts {
  double lrate;     // ln(sqrt(2/3) / (dt * depsp0)): ln(depsp/depsp0) = ln(gamma) + lrate
  double inv_dT;    // 1 / (Tm - T0)
  double gamma_rate;  // depsp > depsp0 <=> gamma > depsp0 dt / sqrt(2/3)
};

__device__ __forceinline__ ReturnConsts return_consts(const double *p, double dt) {
  ReturnConsts c;
  c.lrate = log(kSqrt23 / (dt * p[5]));
  c.inv_dT = 1.0 / (p[6] - p[7]);
  c.gamma_rate = dt * p[5] / kSqrt23;
  return c;
}

// Same yield stress and slope as law_yield_and_slope for the Johnson-Cook
// law, in terms of gamma (depsp = sqrt(2/3) gamma / dt, so
// depsp * dt = sqrt(2/3) gamma) with powers as exp/log pairs and the slope's
// divisions as hardware reciprocals (the slope only sets the Newton step).
__device__ __forceinline__ void jc_yield_and_slope_fast(const double *p, const ReturnConsts &k, double eta_rcp,
                                                        double epsp, double gamma, double T, double &yield,
                                                        double &hard) {
  const double A = p[0], B = p[1], C = p[2], n = p[3], m = p[4], Tm = p[6], T0 = p[7];
  const double lg = log(epsp);
  const double pn = exp(n * lg);
  const double hard0 = A + B * pn;
  const bool rate = gamma > k.gamma_rate;  // depsp > depsp0
  const double visc0 = rate ? 1.0 + C * (log(gamma) + k.lrate) : 1.0;
  double tp1 = 0.0, temp0 = 1.0;
  const bool warm = T > T0;
  const bool soft = warm && T < Tm;
  if (warm) {
    if (T < Tm) {
      tp1 = exp(m * log((T - T0) * k.inv_dT));
      temp0 = 1.0 - tp1;
    } else {
      temp0 = 0.0;
    }
  }
  yield = hard0 * visc0 * temp0;
  double h = n * B * pn * __builtin_amdgcn_rcp(epsp) * visc0 * temp0;
  if (rate) h += hard0 * C * temp0 * __builtin_amdgcn_rcp(kSqrt23 * gamma);
  if (soft) h -= m * eta_rcp * hard0 * hard0 * visc0 * visc0 * temp0 * tp1 * __builtin_amdgcn_rcp(T - T0);
  hard = h;
}

// RampFunction::getValue (RampFunction.C:87-131) with the coefficients of
// RampFunction::setFunction (RampFunction.C:40-84).
__device__ __forceinline__ double ramp_value(int type, double mn, double mx, double x) {
  if (type == DNLG_RAMP_NONE) return 0.0;
  if ((x < mn) || (x > mx)) return 0.0;
  const double mm = mx - mn;
  const double pt = (x - mn) / (mx - mn);
  switch (type) {
    case DNLG_RAMP_LINEAR: return (-2.0 / mm) * pt + (2.0 / mm);
    case DNLG_RAMP_SINUS: return (1.0 / (1.0 - 2.0 / kPI) / mm) * (1.0 - sin((kPI / 2) * pt));
    case DNLG_RAMP_CONSTANT: return 1.0 / mm;
    case DNLG_RAMP_ERF: return exp(-sqd(5.0 * pt)) / (sqrt(kPI) * mm / 10);
    case DNLG_RAMP_SOFT: return (1 / mm) * (kPI * sin(kPI * pt) / 2);
  }
  return 0.0;
}

}  // namespace dnlg
