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

// Logarithms and divisions of the FP64 law evaluation: the device libm and
// IEEE division (the FP64 path, which tracks the reference to rounding) ...
struct LibmOps {
  static constexpr bool kLibm = true;
  __device__ static double lg(double x) { return log(x); }
  __device__ static double ex(double x) { return exp(x); }
  __device__ static double div(double a, double b) { return a / b; }
};

// ... or a shorter logarithm and exponential and a reciprocal from the
// hardware estimate (the FP64 Newton of the mixed path).  log: x = 2^k m
// with m in [1/sqrt(2), sqrt(2)), log m = 2 atanh(s), s = (m - 1) / (m +
// 1), |s| <= 0.1716, by its series to s^21 (truncation below 2.3e-17
// relative), k ln 2 in two parts; within 2 ulp of the exact logarithm on
// [1e-9, 20] and near 1 (30 VALU instructions against the libm's 98).
// Positive normal x only: the law's arguments are.  exp: x = k ln 2 + r,
// |r| <= ln 2 / 2, e^r by its Taylor series to r^13, times 2^k; within 1
// ulp on [-30, 5] (22 instructions against 36).  Reciprocal: the hardware
// estimate refined by two Newton steps.
__device__ __forceinline__ double rcp_f64(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, e, r);
}
// a b + c for a constant c: one multiply-add that reads c from a scalar
// register pair.  Written as a plain fma, the compiler takes the
// two-address form, whose addend is also its destination, and first copies
// the constant into vector registers (two extra instructions per step of a
// polynomial); inside a loop it also keeps every constant of the loop in
// vector registers.
__device__ __forceinline__ double fmak(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}
__device__ __forceinline__ double log_f64_short(double x) {
  double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
  int k = __builtin_amdgcn_frexp_exp(x);
  const bool low = m < 0.70710678118654752;
  m = low ? 2.0 * m : m;
  k = low ? k - 1 : k;
  const double num = m - 1.0, den = m + 1.0;
  const double r = rcp_f64(den);
  double s = num * r;
  s = __builtin_fma(r, __builtin_fma(-den, s, num), s);
  const double z = s * s;
  double q = fmak(2.0 / 21.0, z, 2.0 / 19.0);
  q = fmak(q, z, 2.0 / 17.0);
  q = fmak(q, z, 2.0 / 15.0);
  q = fmak(q, z, 2.0 / 13.0);
  q = fmak(q, z, 2.0 / 11.0);
  q = fmak(q, z, 2.0 / 9.0);
  q = fmak(q, z, 2.0 / 7.0);
  q = fmak(q, z, 2.0 / 5.0);
  q = fmak(q, z, 2.0 / 3.0);
  const double kd = (double)k;
  const double t = __builtin_fma(s * z, q, kd * 1.90821492927058770002e-10);  // + k ln2 (low part)
  return __builtin_fma(kd, 6.93147180369123816490e-01, 2.0 * s + t);          // + k ln2 (high part)
}
__device__ __forceinline__ double exp_f64_short(double x) {
  const double k = __builtin_rint(x * 1.4426950408889634074);  // x / ln 2
  double r = __builtin_fma(-k, 6.93147180369123816490e-01, x);
  r = __builtin_fma(-k, 1.90821492927058770002e-10, r);
  double q = fmak(1.0 / 6227020800.0, r, 1.0 / 479001600.0);  // 1/13!, 1/12!
  q = fmak(q, r, 1.0 / 39916800.0);
  q = fmak(q, r, 1.0 / 3628800.0);
  q = fmak(q, r, 1.0 / 362880.0);
  q = fmak(q, r, 1.0 / 40320.0);
  q = fmak(q, r, 1.0 / 5040.0);
  q = fmak(q, r, 1.0 / 720.0);
  q = fmak(q, r, 1.0 / 120.0);
  q = fmak(q, r, 1.0 / 24.0);
  q = fmak(q, r, 1.0 / 6.0);
  q = __builtin_fma(q, r, 0.5);
  q = __builtin_fma(q, r, 1.0);
  q = __builtin_fma(q, r, 1.0);
  return __builtin_amdgcn_ldexp(q, (int)k);
}
struct ShortOps {
  static constexpr bool kLibm = false;
  __device__ static double lg(double x) { return log_f64_short(x); }
  __device__ static double ex(double x) { return exp_f64_short(x); }
  __device__ static double div(double a, double b) { return a * rcp_f64(b); }
};

// Yield stress and its consistent derivative in one pass: shares the power
// and logarithm terms of JohnsonCookLaw::getYieldStress and
// getDerivateYieldStress (JohnsonCookLaw.C:298-339).  eta_rcp = eta/(rho*Cp).
template <class Ops = LibmOps>
__device__ __forceinline__ void law_yield_and_slope(int law, const double *p, double eta_rcp, double epsp,
                                                    double depsp, double T, double dtime, double &yield,
                                                    double &hard) {
  if (law == DNLG_LAW_JOHNSON_COOK) {
    const double A = p[0], B = p[1], C = p[2], n = p[3], m = p[4], d0 = p[5], Tm = p[6], T0 = p[7];
    const double lg = Ops::lg(epsp);
    const double pn = Ops::ex(n * lg);  // epsp^n
    const double hard0 = A + B * pn;
    const bool rate = depsp > d0;
    const double visc0 = rate ? 1.0 + C * Ops::lg(Ops::div(depsp, d0)) : 1.0;
    double tp1 = 0.0, temp0 = 1.0;
    const bool warm = T > T0;
    if (warm) {
      if (T < Tm) {
        tp1 = Ops::ex(m * Ops::lg(Ops::div(T - T0, Tm - T0)));  // ((T - T0) / (Tm - T0))^m, 0 < base < 1
        temp0 = 1.0 - tp1;
      } else {
        temp0 = 0.0;
      }
    }
    yield = hard0 * visc0 * temp0;
    double h = n * B * Ops::div(pn, epsp) * visc0 * temp0;
    if (rate) h += Ops::div(hard0 * C * temp0, depsp * dtime);
    if (warm && T < Tm) h -= Ops::div(m * eta_rcp * hard0 * hard0 * visc0 * visc0 * temp0 * tp1, T - T0);
    hard = h;
    return;
  }
  if (law == DNLG_LAW_POWER) {
    if constexpr (Ops::kLibm) {
      const double pn = pow(epsp, p[2]);
      yield = p[0] + p[1] * pn;
      hard = p[2] * p[1] * pow(epsp, p[2] - 1.0);
    } else {
      const double pn = Ops::ex(p[2] * Ops::lg(epsp));
      yield = p[0] + p[1] * pn;
      hard = p[2] * p[1] * Ops::div(pn, epsp);
    }
    return;
  }
  const double ex = Ops::ex(-p[2] * epsp);
  yield = p[0] + p[1] * (1.0 - ex) + p[3] * epsp;
  hard = p[1] * p[2] * ex + p[3];
}

// TabularLaw::getYieldStress / getDerivateYieldStress (TabularLaw.C): the
// table's piecewise-linear interpolation and the slope of the segment that
// ends at the first point at or past epsp (DiscreteFunction::getValue /
// getSlope, DiscreteFunction.C:188-216, 284-308).  Returns false outside
// the table (the reference stops there).
__device__ __forceinline__ bool law_tabular(const double *t, int n, double epsp, double &yield, double &hard) {
  const DNLG_GLOBAL double *x = gbl(t), *y = gbl(t) + n;
  if (!(epsp >= x[0] && epsp <= x[n - 1])) return false;
  // the first point i >= 1 with x_i >= epsp, by bisection over the ascending
  // strains (the reference scans the points in order; same segment)
  int lo = 1, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x[mid] >= epsp)
      hi = mid;
    else
      lo = mid + 1;
  }
  const double x0 = x[lo - 1], y0 = y[lo - 1], xi = x[lo], yi = y[lo];
  yield = (xi == epsp) ? yi : ((epsp - x0) / (xi - x0)) * (yi - y0) + y0;
  hard = (yi - y0) / (xi - x0);
  return true;
}

// ANN2SigLaw::getYieldStress / getDerivateYieldStress (ANN2SigLaw.C:173-245)
// on the packed network t (layout of DNLG_LAW_ANN2SIG).  Inputs scaled to
// the training ranges, the rate as log(depsp / logBase) (0 at or below
// logBase, where the slope's rate term divides by logBase instead of
// depsp, as in the reference); output scaled back.  The slope is the
// network's gradient by back-propagation: d/d(ep) + d/d(dep) / dtime +
// eta / (rho Cp) * yield * d/dT.  The first hidden layer's exp(-z) and
// the second's go to per-lane arrays (private memory) of
// DNLG_ANN_MAX_HIDDEN entries.
__device__ __noinline__ void law_ann(const double *t, double epsp, double depsp, double T, double dtime,
                                     double eta_rcp, bool slope, double &yield, double &hard) {
  const DNLG_GLOBAL double *p = gbl(t);
  const int h1 = (int)p[0], h2 = (int)p[1];
  const DNLG_GLOBAL double *w1 = p + 2, *b1 = w1 + 3 * h1, *w2 = b1 + h1, *b2 = w2 + h2 * h1, *w3 = b2 + h2;
  const DNLG_GLOBAL double *tail = w3 + h2;  // b3, logBase, min[4], max[4]
  const double b3 = tail[0], logBase = tail[1];
  const DNLG_GLOBAL double *mn = tail + 2, *mx = tail + 6;
  const double r0 = mx[0] - mn[0], r1 = mx[1] - mn[1], r2 = mx[2] - mn[2], r3 = mx[3] - mn[3];
  double x0 = (epsp - mn[0]) / r0, x1 = 0.0;
  if (depsp > logBase)
    x1 = (log(depsp / logBase) - mn[1]) / r1;
  else
    depsp = logBase;
  const double x2 = (T - mn[2]) / r2;
  double za[DNLG_ANN_MAX_HIDDEN], zc[DNLG_ANN_MAX_HIDDEN];
  for (int i = 0; i < h1; i++) za[i] = exp(-(w1[3 * i] * x0 + w1[3 * i + 1] * x1 + w1[3 * i + 2] * x2 + b1[i]));
  double y = 0.0;
  for (int j = 0; j < h2; j++) {
    double z = 0.0;
    for (int i = 0; i < h1; i++) z += w2[j * h1 + i] * (1.0 / (za[i] + 1.0));
    zc[j] = exp(-(z + b2[j]));
    y += w3[j] * (1.0 / (zc[j] + 1.0));
  }
  y += b3;
  yield = r3 * y + mn[3];
  if (!slope) return;
  // back-propagation: zc becomes dy/d(layer-2 input), zf is dy/d(layer-1 input)
  for (int j = 0; j < h2; j++) {
    const double c1 = zc[j] + 1.0;
    zc[j] = (1.0 / (c1 * c1)) * zc[j] * w3[j];
  }
  double yd0 = 0.0, yd1 = 0.0, yd2 = 0.0;
  for (int i = 0; i < h1; i++) {
    double g = 0.0;
    for (int j = 0; j < h2; j++) g += w2[j * h1 + i] * zc[j];
    const double b = za[i] + 1.0;
    const double zf = g * ((1.0 / (b * b)) * za[i]);
    yd0 += w1[3 * i] * zf;
    yd1 += w1[3 * i + 1] * zf;
    yd2 += w1[3 * i + 2] * zf;
  }
  hard = r3 / r0 * yd0 + (r3 / r1 * yd1 / depsp) / dtime + eta_rcp * yield * (r3 / r2 * yd2);
}

// The hardening law's yield stress (and slope) at (epsp, depsp, T) for the
// FP64 radial returns (the FP64 element kernel, and the mixed kernel's
// variant for the table laws and the Direct integration): a tabular law
// through its table (out of it: `ok` cleared), the closed-form laws
// through the functions above; ANN: the network law.
template <bool ANN>
__device__ __forceinline__ double law_yield_of(CJobDev *C, int law, const double *lp, double epsp, double depsp,
                                               double T, double dt, bool &ok) {
  if (ANN) {
    double y, h;
    law_ann(C->ltab, epsp, depsp, T, dt, 0.0, false, y, h);
    return y;
  }
  if (law == DNLG_LAW_TABULAR) {
    double y = 0.0, h;
    ok = law_tabular(C->ltab, C->n_ltab, epsp, y, h) && ok;
    return y;
  }
  return law_yield(law, lp, epsp, depsp, T);
}
template <bool ANN>
__device__ __forceinline__ void law_eval(CJobDev *C, int law, const double *lp, double eta_rcp, double epsp,
                                         double depsp, double T, double dt, double &yield, double &hard, bool &ok) {
  if (ANN) {
    law_ann(C->ltab, epsp, depsp, T, dt, eta_rcp, true, yield, hard);
    return;
  }
  if (law == DNLG_LAW_TABULAR) {
    yield = hard = 0.0;
    ok = law_tabular(C->ltab, C->n_ltab, epsp, yield, hard) && ok;
    return;
  }
  law_yield_and_slope(law, lp, eta_rcp, epsp, depsp, T, dt, yield, hard);
}

// FP32 law evaluation of the mixed path's radial return.  Same yield stress
// and slope as law_yield_and_slope, written in terms of gamma (depsp =
// sqrt(2/3) gamma / dt, so depsp * dt = sqrt(2/3) gamma), with powers and
// logarithms from the hardware log2 / exp2 (about 1 ulp).  The FP32 residual
// of the return equation carries an error of order 1e-7 x the trial stress,
// i.e. a few 1e-10 in gamma, below the 1e-8 Newton tolerance.
struct LawF {
  int law;
  float p[8];
  float lrate;       // ln(sqrt(2/3) / (dt depsp0)): ln(depsp / depsp0) = ln(gamma) + lrate
  float inv_dT;      // 1 / (Tm - T0)
  float gamma_rate;  // depsp > depsp0  <=>  gamma > depsp0 dt / sqrt(2/3)
  float eta_rcp;     // eta / (rho Cp)
};

// pf: the law parameters in FP32, inv_dT = 1 / (Tm - T0) (Johnson-Cook);
// the rate terms from the cycle's step in FP32 (hardware log2).
__device__ __forceinline__ LawF law_f32_setup(int law, const float (&pf)[8], float inv_dT, float dt, float eta_rcp) {
  LawF L;
  L.law = law;
#pragma unroll
  for (int k = 0; k < 8; k++) L.p[k] = pf[k];
  L.lrate = 0.f;
  L.inv_dT = 0.f;
  L.gamma_rate = 0.f;
  if (law == DNLG_LAW_JOHNSON_COOK) {
    const float r0 = dt * pf[5] * (float)kSqrt32;  // depsp0 dt / sqrt(2/3)
    L.gamma_rate = r0;
    L.lrate = -__builtin_amdgcn_logf(r0) * 0.693147180559945309f;
    L.inv_dT = inv_dT;
  }
  L.eta_rcp = eta_rcp;
  return L;
}

__device__ __forceinline__ float pow_f32(float x, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x)); }

__device__ __forceinline__ void law_f32(const LawF &L, float epsp, float gamma, float T, float &yield,
                                        float &hard) {
  constexpr float kLn2 = 0.693147180559945309f;
  constexpr float s23 = (float)kSqrt23;
  if (L.law == DNLG_LAW_JOHNSON_COOK) {
    const float A = L.p[0], B = L.p[1], C = L.p[2], n = L.p[3], m = L.p[4], Tm = L.p[6], T0 = L.p[7];
    const float pn = pow_f32(epsp, n);
    const float hard0 = A + B * pn;
    const bool rate = gamma > L.gamma_rate;
    const float visc0 = rate ? 1.f + C * (__builtin_amdgcn_logf(gamma) * kLn2 + L.lrate) : 1.f;
    float tp1 = 0.f, temp0 = 1.f;
    const bool warm = T > T0;
    const bool soft = warm && T < Tm;
    if (warm) {
      if (T < Tm) {
        tp1 = pow_f32((T - T0) * L.inv_dT, m);
        temp0 = 1.f - tp1;
      } else {
        temp0 = 0.f;
      }
    }
    yield = hard0 * visc0 * temp0;
    float h = n * B * pn * __builtin_amdgcn_rcpf(epsp) * visc0 * temp0;
    if (rate) h += hard0 * C * temp0 * __builtin_amdgcn_rcpf(s23 * gamma);
    if (soft) h -= m * L.eta_rcp * hard0 * hard0 * visc0 * visc0 * temp0 * tp1 * __builtin_amdgcn_rcpf(T - T0);
    hard = h;
    return;
  }
  if (L.law == DNLG_LAW_POWER) {
    const float pn = pow_f32(epsp, L.p[2]);
    yield = L.p[0] + L.p[1] * pn;
    hard = L.p[2] * L.p[1] * pn * __builtin_amdgcn_rcpf(epsp);
    return;
  }
  const float ex = __builtin_amdgcn_exp2f(-L.p[2] * epsp * 1.44269504088896341f);
  yield = L.p[0] + L.p[1] * (1.f - ex) + L.p[3] * epsp;
  hard = L.p[1] * L.p[2] * ex + L.p[3];
}

// RampFunction::getValue (RampFunction.C:87-131) with the coefficients of
// RampFunction::setFunction (RampFunction.C:40-84).
__device__ __noinline__ double ramp_value(int type, double mn, double mx, double x) {
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
