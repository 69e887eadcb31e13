// elem_mixed.h -- mixed-precision element kernel (instantiated by elem_mixed.hip).
//
// Same phases as k_elem_fp64 (Element.C:1103-1161, 362-404, 407-622,
// 1243-1256) with FP32 geometry, strain, stress, radial return and forces,
// and FP64 for the nodal kinematics and the accumulated equivalent plastic
// strain, temperature and energies.  Divisions and square roots of the FP32
// phases use the hardware reciprocal and square root (about 1 ulp).
//
// Strain increment in increment form: with H = d(du)/dx the right
// Cauchy-Green tensor is C = I + X, X = H + H^T + H^T H (small), and
//   ln U = 1/2 ln(I + X)   = 1/2 sum_k (-1)^(k+1) X^k / k,
//   U^-1 = (I + X)^(-1/2)  = sum_k binom(-1/2, k) X^k,
//   R = F U^-1 = (I + H) U^-1,
// evaluated with powers of X up to X^6 (X^4 when every lane of the wave has
// ||X||_F <= 1e-2, X^3 below 1e-3); the truncation stays under 2e-9
// relative, i.e. under FP32 rounding.  Larger increments (> 0.05) take an
// FP64 eigen-decomposition path (polar_lnU_jacobi).  This avoids the
// catastrophic cancellation of forming F^T F - I and ln(lambda ~ 1) in FP32.
//
// Element geometry uses the tensor-product structure of the 8-node hex
// (4-node quad): with s_d(b) = +-1 the local coordinate signs of node b, the
// Walsh modes M_q = sum_b prod_{d in q} s_d(b) x_b give
//   x(xi) = 2^-D sum_q M_q prod_{d in q} xi_d,
// so the Jacobian at a Gauss point needs 3 FMAs per entry.  The modes are a
// D-stage butterfly over the element's lanes (lane b holds node b).  Forces
// use the transpose: with T_p = w_p J_p^-1 sigma_p per Gauss point,
//   f_a = -sum_p sum_k dN_a/dxi_k(p) T_p,k
// and sum_p dN_a/dxi_k(p) T_p,k is a short combination of the Walsh
// transform of T over the Gauss-point lanes.  Lane b <-> node / Gauss point
// b ^ ((b >> 1) & 1) of the reference numbering (ElHex8N3D.C:30-231,
// ElQua4NAx.C:30-110), whose local signs are the bits of b.

#pragma once

#include "control.h"
#include "material.h"
#include "polar_fp64.h"

namespace dnlg {

// Symmetric tensors here are [xx, xy, xz, yy, yz, zz].  B2: the 2+1 block
// structure of the plane and axisymmetric elements (xz = yz = 0, and in the
// velocity gradient H02 = H12 = H20 = H21 = 0).  The B2 forms keep the terms
// of the general forms that remain, in the same order and grouping, so their
// nonzero results carry the same bits; the dropped terms are exact zeros.

// C = A B for commuting symmetric A, B (powers of one matrix): symmetric
template <bool B2>
__device__ __forceinline__ void sym_mul(const float A[6], const float B[6], float C[6]) {
  if constexpr (B2) {
    C[0] = A[0] * B[0] + A[1] * B[1];
    C[1] = A[0] * B[1] + A[1] * B[3];
    C[2] = 0.f;
    C[3] = A[1] * B[1] + A[3] * B[3];
    C[4] = 0.f;
    C[5] = A[5] * B[5];
  } else {
    C[0] = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
    C[1] = A[0] * B[1] + A[1] * B[3] + A[2] * B[4];
    C[2] = A[0] * B[2] + A[1] * B[4] + A[2] * B[5];
    C[3] = A[1] * B[1] + A[3] * B[3] + A[4] * B[4];
    C[4] = A[1] * B[2] + A[3] * B[4] + A[4] * B[5];
    C[5] = A[2] * B[2] + A[4] * B[4] + A[5] * B[5];
  }
}

// x : y for symmetric x, y
template <bool B2>
__device__ __forceinline__ float sym_dot(const float x[6], const float y[6]) {
  if constexpr (B2)
    return x[0] * y[0] + x[3] * y[3] + x[5] * y[5] + 2.f * (x[1] * y[1]);
  else
    return x[0] * y[0] + x[3] * y[3] + x[5] * y[5] + 2.f * (x[1] * y[1] + x[2] * y[2] + x[4] * y[4]);
}

// R S R^T for symmetric S as (R S) R^T: 45 FMAs (SymTensor2::dotRxRT,
// SymTensor2.C:512-525, which the FP64 path follows term by term); B2: the
// block form (R02 = R12 = R20 = R21 = 0), 14 FMAs.
template <bool B2>
__device__ __forceinline__ void rotate_sym_f32(const float r[9], float S[6]) {
  if constexpr (B2) {
    const float m00 = r[0] * S[0] + r[1] * S[1], m01 = r[0] * S[1] + r[1] * S[3];
    const float m10 = r[3] * S[0] + r[4] * S[1], m11 = r[3] * S[1] + r[4] * S[3];
    const float m22 = r[8] * S[5];
    S[0] = m00 * r[0] + m01 * r[1];
    S[1] = m00 * r[3] + m01 * r[4];
    S[2] = 0.f;
    S[3] = m10 * r[3] + m11 * r[4];
    S[4] = 0.f;
    S[5] = m22 * r[8];
    return;
  }
  const float s3[3][3] = {{S[0], S[1], S[2]}, {S[1], S[3], S[4]}, {S[2], S[4], S[5]}};
  float m[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) m[i][j] = r[3 * i] * s3[0][j] + r[3 * i + 1] * s3[1][j] + r[3 * i + 2] * s3[2][j];
  S[0] = m[0][0] * r[0] + m[0][1] * r[1] + m[0][2] * r[2];
  S[1] = m[0][0] * r[3] + m[0][1] * r[4] + m[0][2] * r[5];
  S[2] = m[0][0] * r[6] + m[0][1] * r[7] + m[0][2] * r[8];
  S[3] = m[1][0] * r[3] + m[1][1] * r[4] + m[1][2] * r[5];
  S[4] = m[1][0] * r[6] + m[1][1] * r[7] + m[1][2] * r[8];
  S[5] = m[2][0] * r[6] + m[2][1] * r[7] + m[2][2] * r[8];
}

// Large increments: FP64 Jacobi eigen path (rarely taken).
__device__ __forceinline__ void polar_lnU_large(const float *H, float *dE, float *R) {
  double F[9], e64[6], r64[9];
#pragma unroll
  for (int k = 0; k < 9; k++) F[k] = (double)H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
  polar_lnU_jacobi(F, e64, r64);
#pragma unroll
  for (int c = 0; c < 6; c++) dE[c] = (float)e64[c];
#pragma unroll
  for (int c = 0; c < 9; c++) R[c] = (float)r64[c];
}

// H (row-major 3x3, FP32) -> dE = ln U (sym), R = F U^-1 (row-major).
// B2: H has the 2+1 block structure; Z8 (plane strain): also H22 = 0, so
// ln U and R keep the zz entries 0 and 1.
template <bool B2, bool Z8>
__device__ __forceinline__ void polar_lnU_mixed(const float H[9], float dE[6], float R[9], int order) {
  float X[6];
  if constexpr (B2) {
    X[0] = 2.f * H[0] + (H[0] * H[0] + H[3] * H[3]);
    X[1] = H[1] + H[3] + (H[0] * H[1] + H[3] * H[4]);
    X[2] = 0.f;
    X[3] = 2.f * H[4] + (H[1] * H[1] + H[4] * H[4]);
    X[4] = 0.f;
    X[5] = Z8 ? 0.f : 2.f * H[8] + (H[8] * H[8]);
  } else {
    X[0] = 2.f * H[0] + (H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
    X[1] = H[1] + H[3] + (H[0] * H[1] + H[3] * H[4] + H[6] * H[7]);
    X[2] = H[2] + H[6] + (H[0] * H[2] + H[3] * H[5] + H[6] * H[8]);
    X[3] = 2.f * H[4] + (H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
    X[4] = H[5] + H[7] + (H[1] * H[2] + H[4] * H[5] + H[7] * H[8]);
    X[5] = 2.f * H[8] + (H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
  }
  const float nrm2 = sym_dot<B2>(X, X);
  if (nrm2 > 0.05f * 0.05f) {
    polar_lnU_large(H, dE, R);
  } else {
    // powers to X^order (order 3, 4 or 6, uniform over the wave, so the
    // branches below do not diverge); each order's sums are the leading
    // terms of the order-6 sums in the same order, whose remaining terms
    // are exact zeros at the lower orders
    float X2[6], X3[6], Ui[6];
    sym_mul<B2>(X, X, X2);
    sym_mul<B2>(X2, X, X3);
    if (order >= 6) {
      float X4[6], X5[6], X6[6];
      sym_mul<B2>(X2, X2, X4);
      sym_mul<B2>(X4, X, X5);
      sym_mul<B2>(X3, X3, X6);
#pragma unroll
      for (int c = 0; c < 6; c++) {
        dE[c] = 0.5f * (X[c] - X2[c] * (1.f / 2.f) + X3[c] * (1.f / 3.f) - X4[c] * (1.f / 4.f) +
                        X5[c] * (1.f / 5.f) - X6[c] * (1.f / 6.f));
        Ui[c] = -X[c] * (1.f / 2.f) + X2[c] * (3.f / 8.f) - X3[c] * (5.f / 16.f) + X4[c] * (35.f / 128.f) -
                X5[c] * (63.f / 256.f) + X6[c] * (231.f / 1024.f);
      }
    } else if (order >= 4) {
      float X4[6];
      sym_mul<B2>(X2, X2, X4);
#pragma unroll
      for (int c = 0; c < 6; c++) {
        dE[c] = 0.5f * (X[c] - X2[c] * (1.f / 2.f) + X3[c] * (1.f / 3.f) - X4[c] * (1.f / 4.f));
        Ui[c] = -X[c] * (1.f / 2.f) + X2[c] * (3.f / 8.f) - X3[c] * (5.f / 16.f) + X4[c] * (35.f / 128.f);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 6; c++) {
        dE[c] = 0.5f * (X[c] - X2[c] * (1.f / 2.f) + X3[c] * (1.f / 3.f));
        Ui[c] = -X[c] * (1.f / 2.f) + X2[c] * (3.f / 8.f) - X3[c] * (5.f / 16.f);
      }
    }
    Ui[0] += 1.f;
    Ui[3] += 1.f;
    Ui[5] += 1.f;
    // R = (I + H) Ui
    if constexpr (B2) {
      R[0] = Ui[0] + (H[0] * Ui[0] + H[1] * Ui[1]);
      R[1] = Ui[1] + (H[0] * Ui[1] + H[1] * Ui[3]);
      R[3] = Ui[1] + (H[3] * Ui[0] + H[4] * Ui[1]);
      R[4] = Ui[3] + (H[3] * Ui[1] + H[4] * Ui[3]);
      const float t = H[8] * Ui[5];  // rounded on its own, as in the general form
      R[8] = Ui[5] + t;
    } else {
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const float a0 = H[3 * r], a1 = H[3 * r + 1], a2 = H[3 * r + 2];
        const float u0 = (r == 0) ? Ui[0] : (r == 1 ? Ui[1] : Ui[2]);
        const float u1 = (r == 0) ? Ui[1] : (r == 1 ? Ui[3] : Ui[4]);
        const float u2 = (r == 0) ? Ui[2] : (r == 1 ? Ui[4] : Ui[5]);
        R[3 * r + 0] = u0 + (a0 * Ui[0] + a1 * Ui[1] + a2 * Ui[2]);
        R[3 * r + 1] = u1 + (a0 * Ui[1] + a1 * Ui[3] + a2 * Ui[4]);
        R[3 * r + 2] = u2 + (a0 * Ui[2] + a1 * Ui[4] + a2 * Ui[5]);
      }
    }
  }
  if constexpr (B2) {
    dE[2] = dE[4] = 0.f;
    R[2] = R[5] = R[6] = R[7] = 0.f;
    if constexpr (Z8) {
      dE[5] = 0.f;
      R[8] = 1.f;
    }
  }
}


// Walsh-Hadamard transform over the NB-bit lane group of an element: lane q
// ends with sum_b prod_{d in q} s_d(b) v_b.
// Stage d pairs the lanes differing in bit d: the lane with the bit clear
// keeps v_lo + v_hi, the one with the bit set v_hi - v_lo, i.e. one FMA
// v + cs[d] * partner with cs[d] = -1 where bit d is set, +1 otherwise.
// Value of lane (this lane ^ 2^d), d < 3: d = 0, 1 by DPP quad permutations
// [1,0,3,2] / [2,3,0,1] (one VALU instruction each); d = 2 by a swizzle in
// bit-mask mode (and 0x1F, xor 4), which runs in the LDS pipeline without
// touching LDS memory -- in DPP it took a copy and two row shifts under bank
// masks (row_shl:4 for banks 0 and 2, row_shr:4 for 1 and 3), three VALU
// instructions in a kernel bound by its VALU issue.
template <int D>
__device__ __forceinline__ float lane_xor(float v) {
  const int x = __float_as_int(v);
  if constexpr (D == 0) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true));
  if constexpr (D == 1) return __int_as_float(__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true));
  return __int_as_float(__builtin_amdgcn_ds_swizzle(x, (4 << 10) | 0x1F));
}

template <int NB>
__device__ __forceinline__ float wht(float v, const float (&cs)[3]) {
  v = fmaf(cs[0], lane_xor<0>(v), v);
  v = fmaf(cs[1], lane_xor<1>(v), v);
  if constexpr (NB == 3) v = fmaf(cs[2], lane_xor<2>(v), v);
  return v;
}

template <int NB>
__device__ __forceinline__ float group_sum(float v) {
  v += lane_xor<0>(v);
  v += lane_xor<1>(v);
  if constexpr (NB == 3) v += lane_xor<2>(v);
  return v;
}

__device__ __forceinline__ float sgn(int b, int d) { return ((b >> d) & 1) ? 1.f : -1.f; }

// Jacobian J_ij = dx_i/dxi_j at the Gauss point xi from the Walsh modes mx
// of the element's coordinates (x(xi) = sum_q mx_q prod_{d in q} xi_d),
// its determinant and inverse.
template <int DIM, int NN>
__device__ __forceinline__ void jacobian_inverse(const float (&mx)[NN][DIM], const float (&xi)[DIM], float &detJ,
                                                 float (&inv)[DIM][DIM]) {
  float Jm[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; i++) {
    if constexpr (DIM == 3) {
      Jm[i][0] = mx[1][i] + mx[3][i] * xi[1] + mx[5][i] * xi[2] + mx[7][i] * (xi[1] * xi[2]);
      Jm[i][1] = mx[2][i] + mx[3][i] * xi[0] + mx[6][i] * xi[2] + mx[7][i] * (xi[0] * xi[2]);
      Jm[i][2] = mx[4][i] + mx[5][i] * xi[0] + mx[6][i] * xi[1] + mx[7][i] * (xi[0] * xi[1]);
    } else {
      Jm[i][0] = mx[1][i] + mx[3][i] * xi[1];
      Jm[i][1] = mx[2][i] + mx[3][i] * xi[0];
    }
  }
  if constexpr (DIM == 3) {
    const float *dd = &Jm[0][0];
    detJ = dd[1] * dd[5] * dd[6] - dd[2] * dd[4] * dd[6] + dd[2] * dd[3] * dd[7] - dd[0] * dd[5] * dd[7] -
           dd[1] * dd[3] * dd[8] + dd[0] * dd[4] * dd[8];
    const float id = __builtin_amdgcn_rcpf(detJ);
    inv[0][0] = (dd[4] * dd[8] - dd[5] * dd[7]) * id;
    inv[0][1] = (dd[2] * dd[7] - dd[1] * dd[8]) * id;
    inv[0][2] = (dd[1] * dd[5] - dd[2] * dd[4]) * id;
    inv[1][0] = (dd[5] * dd[6] - dd[3] * dd[8]) * id;
    inv[1][1] = (dd[0] * dd[8] - dd[2] * dd[6]) * id;
    inv[1][2] = (dd[2] * dd[3] - dd[0] * dd[5]) * id;
    inv[2][0] = (dd[3] * dd[7] - dd[4] * dd[6]) * id;
    inv[2][1] = (dd[1] * dd[6] - dd[0] * dd[7]) * id;
    inv[2][2] = (dd[0] * dd[4] - dd[1] * dd[3]) * id;
  } else {
    detJ = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
    const float id = __builtin_amdgcn_rcpf(detJ);
    inv[0][0] = Jm[1][1] * id;
    inv[0][1] = -Jm[0][1] * id;
    inv[1][0] = -Jm[1][0] * id;
    inv[1][1] = Jm[0][0] * id;
  }
}

// Per-lane rows of R float4 in LDS, exchanged inside a wave (the NN lanes of
// an element are in one wave; one wave's LDS accesses execute in order, and
// wave_barrier keeps the compiler from moving them across the exchange).
// Each element's rows are followed by one float4 of padding, so that the
// element groups of a wave, which read one row each, fall on different banks.
template <int NN, int R>
struct LaneRows {
  static constexpr int kGroup = NN * R + 1;  // float4 per element
  __device__ static int at(int le, int lane, int v) { return le * kGroup + lane * R + v; }
};

// Newton on gamma for Element::computeStress's radial return
// (Element.C:485-601) in FP32: the bisection safeguard of the reference, the
// law from law_f32.  Stops at the reference tolerance, or 4 ulp of gamma
// where FP32 cannot resolve it; after a Newton step below 2e-4 gamma the
// remaining error is of order |f''/f'| dg^2 < 1e-11 (hardening, rate and
// thermal terms of the laws here), so the confirming iteration is left out.
// In: the trial stress Strial = sqrt(3/2) |s_trial|, yield = the stored
// yield stress, g = the previous gamma (the start unless ep0 == 0, as in the
// reference).  Out: gamma and the yield stress of the last iterate, and in
// `last` the state an FP64 return resumes at -- a gamma, the yield stress
// its temperature is formed with and the bracket: the start of the last
// iteration, or, when that iteration ended on a Newton step of at least the
// reference's tolerance 1e-8, the start of the next one (the reference takes
// that iterate too, and a Newton step from it lands on the root to
// rounding).  Returns false at the iteration limit.
struct NewtonState {
  float g, y, gmin, gmax;
};
__device__ __forceinline__ bool radial_return_fp32(const LawF &L, int max_it, float Strial, float TwoG32, float heat,
                                                   float Tc, double ep0, float Snorm0, float &yield, float &g,
                                                   NewtonState &last) {
  constexpr float s23 = (float)kSqrt23;
  const float ep0f = (float)ep0;
  float gmin = 0.f, gmax = (Strial - yield) * __builtin_amdgcn_rcpf(TwoG32);
  if (ep0 == 0.0) g = (float)(kSqrt32 * kGammaInitial);
  float epsp = ep0f + s23 * g;
  float T = Tc + 0.5f * g * heat * (s23 * yield + Snorm0);
  int iterate = 0;
  while (true) {
    last = NewtonState{g, yield, gmin, gmax};
    float hard;
    law_f32(L, epsp, g, T, yield, hard);
    const float fun = Strial - g * TwoG32 - yield;
    if (fun < 0.f)
      gmax = g;
    else
      gmin = g;
    const float dfun = TwoG32 + s23 * hard;
    float dg = fun * __builtin_amdgcn_rcpf(dfun);
    g += dg;
    bool newton = true;
    if ((gmax - g) * (g - gmin) < 0.f) {
      dg = 0.5f * (gmax - gmin);
      g = gmin + dg;
      newton = false;
    }
    if (fabsf(dg) < fmaxf((float)kTolNR, (newton ? 2e-4f : 4.8e-7f) * g)) {
      if (newton && fabsf(dg) >= (float)kTolNR) last = NewtonState{g, yield, gmin, gmax};
      return true;
    }
    epsp = ep0f + s23 * g;
    T = Tc + 0.5f * g * heat * (s23 * yield + Snorm0);
    if (++iterate > max_it) return false;
  }
}

// The same Newton in FP64 with the reference's stopping rule (|dgamma| <
// 1e-8, at most 250 iterations) and operation order, the yield stress and
// slope from law_yield_and_slope in FP64 (logarithms, exponentials, reciprocals
// within 2 ulp, ShortOps), resuming the FP32 iteration where
// radial_return_fp32 leaves it: gamma g, the yield stress y that
// iteration's temperature is formed with, and the bracket [gmin, gmax].
// From there the reference's own iterates are taken again in FP64: its path
// is reproduced up to the FP32 rounding of the resumed state, which the
// Newton step contracts, so one FP64 iteration ends where the reference
// ends (more after a bisection, or where the two precisions fall on
// different sides of the tolerance).  Resuming further back would only
// cost iterations; resuming anywhere else would not do: the reference's
// result depends on its path -- its absolute tolerance
// leaves a small gamma (first yield) 1e-3 to 1e-1 short of the root, and
// when the root lies above the initial bound (Strial - y_stored) /
// (sqrt(3/2) 2G) a start above the bound moves gmax up to itself and
// reaches the root while a start below it is held under the bound.  Out:
// gamma and the FP64 yield stress of its last iterate.  Returns false at the
// iteration limit.
// rdt = 1 / dt (the rate is sqrt(2/3) gamma rdt).
// law_at(epsp, depsp, T, yield, slope) evaluates the law; it returns false
// where the law has no value (outside a table), which ends the iteration
// (the caller reports it).
template <class Law>
__device__ __forceinline__ bool radial_return_fp64(const Law &law_at, int max_it, double heatFrac, double TwoG32,
                                                   double Strial, double Snorm0, double Tcur, double ep0,
                                                   double rdt, double gmin, double gmax, double &y, double &g) {
  double depsp = kSqrt23 * g * rdt;
  double epsp = ep0 + kSqrt23 * g;
  double T = Tcur + 0.5 * g * heatFrac * (kSqrt23 * y + Snorm0);
  int iterate = 0;
  while (true) {
    double hard;
    if (!law_at(epsp, depsp, T, y, hard)) return true;
    const double fun = Strial - g * TwoG32 - y;
    if (fun < 0.0)
      gmax = g;
    else
      gmin = g;
    const double dfun = TwoG32 + kSqrt23 * hard;
    double dg = fun * rcp_f64(dfun);
    g += dg;
    if ((gmax - g) * (g - gmin) < 0.0) {
      dg = 0.5 * (gmax - gmin);
      g = gmin + dg;
    }
    if (fabs(dg) < kTolNR) return true;
    depsp = kSqrt23 * g * rdt;
    epsp = ep0 + kSqrt23 * g;
    T = Tcur + 0.5 * g * heatFrac * (kSqrt23 * y + Snorm0);
    if (++iterate > max_it) return false;
  }
}

// RR64: the radial return in FP64 (DNLG_PREC_MIXED); otherwise in FP32
// (DNLG_PREC_MIXED_RR32).
// WAVES > 0: that many waves per SIMD requested from the compiler (registers
// capped, spills allowed); 0: none requested.
// XL (with RR64): the variant for the laws that have no FP32 form -- the
// tabular and network laws -- and for the Direct integration: the FP64
// return of the FP64 path (Element.C:485-601, 625-768) from the
// reference's own starting state, on the FP32 trial stress.
template <int NN, int DIM, int BLOCK, bool AXI, bool RR64, int WAVES, bool XL = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1))) void k_elem_mixed(
    JobDev *const *jobs, const Task *tasks, int ctl, int derive) {
  constexpr int NB = DIM;               // lanes per element = 2^DIM = NN
  constexpr float inv2D = 1.f / NN;     // 2^-D
  // plane and axisymmetric quads: 2+1 block structure, no xz / yz stress
  // (those entries of the stored stress stay 0); plane strain: no hoop strain
  constexpr bool B2 = DIM == 2, Z8 = B2 && !AXI;
  const float g = 0.577350269189625764f;  // Gauss abscissa 1/sqrt(3)

  // raised issue priority until this wave's element modes are in its
  // registers (below): a wave that starts while others compute gets its
  // loads in flight and its butterflies and mode exchange done first
  __builtin_amdgcn_s_setprio(2);
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  // Every job-record field of the control derivation and of the loads below
  // in one batch of scalar loads ahead of the first branch (the empty asm
  // takes them as operands, so none is sunk behind a branch to cost a round
  // trip of its own; cycle_control reads the same constant-space words).
  const int64_t N = C->N, G = C->G;
  const DNLG_GLOBAL int32_t *conn = gbl(C->conn), *e2slot = gbl(C->e2slot);
  const DNLG_GLOBAL double *xg = gbl(C->x), *fu0 = gbl(C->fu[0]), *fu1 = gbl(C->fu[1]);
  const DNLG_GLOBAL float *sig32 = gbl(C->sig32), *sy32 = gbl(C->sy32), *gam32 = gbl(C->gam32);
  const DNLG_GLOBAL double *Tg = gbl(C->T), *epg = gbl(C->ep);
  {
    const Control c0 = load_control(C, ctl);
    asm volatile("" ::"s"(tk.first), "s"(tk.end), "s"(tk.blk), "s"(c0.t), "s"(c0.dt), "s"(c0.up_to), "s"(c0.inc),
                 "s"(c0.f0), "s"(c0.running), "s"(c0.pending), "s"(C->sync.st[0]), "s"(C->sync.st[1]),
                 "s"(C->pi.stall), "s"(C->res.hold), "s"(C->dt_freq), "s"(C->density_always), "s"(N), "s"(G),
                 "s"(conn), "s"(e2slot), "s"(xg), "s"(fu0), "s"(fu1), "s"(sig32), "s"(sy32), "s"(gam32), "s"(Tg),
                 "s"(epg));
  }
  const Control cc = cycle_control(J, C, ctl, derive, tk.blk);
  if (!cc.running) return;
  const int par = (int)(cc.inc & 1);
  const int lane = threadIdx.x & 63;
  const int b = lane & (NN - 1);            // lane within the element group
  const int base = lane & ~(NN - 1);        // first lane of the group
  const int le = threadIdx.x / NN;
  const int64_t e = tk.first + le;
  const bool valid = e < tk.end;

  const int ref = b ^ ((b >> 1) & 1);       // node / IP index in the reference order
  const int f0 = cc.f0;

  // ---- loads: the element's connectivity, then this lane's IP state (it
  // does not depend on the connectivity, so its latency overlaps the node
  // loads and the geometry), then node b.  Idle lanes of a partial block
  // read the block's first element (in range) and discard the values, so the
  // loads need no branch, which would make the hardware wait for them there.
  // Per-IP arrays are addressed from the block's first IP (uniform) with a
  // lane-local 32-bit byte offset: ip = this lane's IP within the block.
  const int64_t g0 = tk.first * NN;
  const uint32_t ip = (uint32_t)((valid ? le * NN : 0) + ref), e0 = (uint32_t)(valid ? le * NN : 0);
  const int nd = at(conn + g0, 4 * ip);
  const int n0 = at(conn + g0, 4 * e0);
  const int slot = at(e2slot + g0, 4 * ip);  // incidence slot of this lane's node, where its force goes
  float sig0[6];                             // stress, yield stress, gamma of the last cycle
#pragma unroll
  for (int c = 0; c < 6; c++) sig0[c] = (B2 && (c == 2 || c == 4)) ? 0.f : at(sig32 + c * G + g0, 4 * ip);
  const float sy0 = at(sy32 + g0, 4 * ip), gam0 = at(gam32 + g0, 4 * ip);
  const double Tcur = at(Tg + g0, 8 * ip), ep0 = at(epg + g0, 8 * ip);
  // node b of the element: coordinates relative to node 0, increment
  const DNLG_GLOBAL double *u0 = f0 ? fu1 : fu0;
  float xl[DIM], ul[DIM];
  double x0r = 0.0;
  {
    double xb[DIM], xa[DIM], ub[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      xb[d] = at(xg + d * N, 8 * (uint32_t)nd);
      xa[d] = at(xg + d * N, 8 * (uint32_t)n0);
      ub[d] = at(u0 + d * N, 8 * (uint32_t)nd);
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      // idle lanes: a dummy unit element keeps them finite
      xl[d] = valid ? (float)(xb[d] - xa[d]) : sgn(b, d);
      ul[d] = valid ? (float)ub[d] : 0.f;
    }
    if constexpr (AXI) x0r = valid ? xa[0] : 0.0;
  }

  // ---- Walsh modes of coordinates and increments, scaled by 2^-D
  float cs[3];
#pragma unroll
  for (int d = 0; d < 3; d++) cs[d] = ((b >> d) & 1) ? -1.f : 1.f;
  float Mx[DIM], Mu[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    Mx[d] = inv2D * wht<NB>(xl[d], cs);
    Mu[d] = inv2D * wht<NB>(ul[d], cs);
  }
  // LDS, one region per wave (the waves of a block are not synchronised):
  // the inverse Jacobian and det J of each lane, kept from the geometry to
  // the forces, then one area that the three phases take in turn -- the
  // element modes (geometry), the values held across the radial return,
  // the transformed force terms (forces).  A wave's LDS accesses execute in
  // order, so a phase's writes follow the previous phase's reads.
  constexpr int MR = (2 * DIM + 3) / 4;  // float4 per lane row of modes
  constexpr int FR = DIM == 3 ? 3 : 2;   // float4 per lane row of force terms
  typedef LaneRows<NN, MR> ModeRows;
  typedef LaneRows<NN, FR> ForceRows;
  constexpr int EPW = 64 / NN;           // elements per wave
  constexpr int NJ = DIM * DIM + 1;      // J^-1 entries and det J
  constexpr int kModeF4 = EPW * ModeRows::kGroup, kKeepF4 = 9 * 16, kWtF4 = EPW * ForceRows::kGroup;
  constexpr int kPhaseF4 = kModeF4 > kKeepF4 ? (kModeF4 > kWtF4 ? kModeF4 : kWtF4) : (kKeepF4 > kWtF4 ? kKeepF4 : kWtF4);
  constexpr int kWaveF4 = NJ * 16 + kPhaseF4;
  __shared__ float4 s_lds[(BLOCK / 64) * kWaveF4];
  float4 *const s_wave = s_lds + (threadIdx.x >> 6) * kWaveF4;
  float *const s_jinv = (float *)s_wave;               // [NJ][64]
  float4 *const s_mode = s_wave + NJ * 16;             // geometry
  float *const s_keep = (float *)(s_wave + NJ * 16);   // radial return: [9][64]
  float4 *const s_wt = s_wave + NJ * 16;               // forces
  const int lw = lane / NN;                            // element within the wave
  // modes of the group to every lane of the element: mx[q][d], mu[q][d]
  {
    float row[4 * MR];
#pragma unroll
    for (int k = 0; k < 4 * MR; k++) row[k] = 0.f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      row[d] = Mx[d];
      row[DIM + d] = Mu[d];
    }
#pragma unroll
    for (int v = 0; v < MR; v++)
      s_mode[ModeRows::at(lw, b, v)] = make_float4(row[4 * v], row[4 * v + 1], row[4 * v + 2], row[4 * v + 3]);
  }
  __builtin_amdgcn_wave_barrier();
  float mx[NN][DIM], mu[NN][DIM];
#pragma unroll
  for (int q = 0; q < NN; q++) {
    float row[4 * MR];
#pragma unroll
    for (int v = 0; v < MR; v++) {
      const float4 t = s_mode[ModeRows::at(lw, q, v)];
      row[4 * v] = t.x;
      row[4 * v + 1] = t.y;
      row[4 * v + 2] = t.z;
      row[4 * v + 3] = t.w;
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      mx[q][d] = row[d];
      mu[q][d] = row[DIM + d];
    }
  }
  __builtin_amdgcn_s_setprio(0);
  // Gauss point of this lane
  float xi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) xi[d] = g * sgn(b, d);

  // ---- Jacobian, its inverse, and G_ij = du_i/dxi_j
  float detJ, inv[DIM][DIM];
  jacobian_inverse<DIM>(mx, xi, detJ, inv);
#pragma unroll
  for (int k = 0; k < DIM; k++)
#pragma unroll
    for (int j = 0; j < DIM; j++) s_jinv[(k * DIM + j) * 64 + lane] = inv[k][j];
  s_jinv[DIM * DIM * 64 + lane] = detJ;
  float Gm[DIM][DIM];
#pragma unroll
  for (int i = 0; i < DIM; i++) {
    if constexpr (DIM == 3) {
      Gm[i][0] = mu[1][i] + mu[3][i] * xi[1] + mu[5][i] * xi[2] + mu[7][i] * (xi[1] * xi[2]);
      Gm[i][1] = mu[2][i] + mu[3][i] * xi[0] + mu[6][i] * xi[2] + mu[7][i] * (xi[0] * xi[2]);
      Gm[i][2] = mu[4][i] + mu[5][i] * xi[0] + mu[6][i] * xi[1] + mu[7][i] * (xi[0] * xi[1]);
    } else {
      Gm[i][0] = mu[1][i] + mu[3][i] * xi[1];
      Gm[i][1] = mu[2][i] + mu[3][i] * xi[0];
    }
  }
  if (valid && detJ < 0.f) set_status(J, DNLG_E_NEG_JACOBIAN, par);
  // H = du/dx = G J^-1 (Element3D.C:172-195, ElementAxi.C:78-100)
  float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < DIM; i++)
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      float h = 0.f;
#pragma unroll
      for (int k = 0; k < DIM; k++) h += Gm[i][k] * inv[k][j];
      H[3 * i + j] = h;
    }
  float radius = 1.f;
  if constexpr (AXI) {
    const float rl = mx[0][0] + mx[1][0] * xi[0] + mx[2][0] * xi[1] + mx[3][0] * (xi[0] * xi[1]);
    const float ur = mu[0][0] + mu[1][0] * xi[0] + mu[2][0] * xi[1] + mu[3][0] * (xi[0] * xi[1]);
    radius = (float)(x0r + (double)rl);
    H[8] = ur * __builtin_amdgcn_rcpf(radius);
  }
  // series order, uniform over the wave, from the linear part of X
  float Xn2;
  {
    const float a0 = 2.f * H[0], a3 = 2.f * H[4], a5 = 2.f * H[8];
    const float a1 = H[1] + H[3];
    if constexpr (B2) {
      Xn2 = a0 * a0 + a3 * a3 + a5 * a5 + 2.f * (a1 * a1);
    } else {
      const float a2 = H[2] + H[6], a4 = H[5] + H[7];
      Xn2 = a0 * a0 + a3 * a3 + a5 * a5 + 2.f * (a1 * a1 + a2 * a2 + a4 * a4);
    }
  }
  const int order = __any(Xn2 > 1e-4f) ? 6 : (__any(Xn2 > 1e-6f) ? 4 : 3);
  float dE[6], R[9];
  polar_lnU_mixed<B2, Z8>(H, dE, R, order);

  // ---- pressure: element mean of tr(dE) (Element.C:1125-1143)
  const float inc = group_sum<NB>(dE[0] + dE[3] + dE[5]) * inv2D;

  float T9[DIM][DIM];
  float hoop = 0.f;
  float sig[6];
  if (valid) {
    float pressure = (sig0[0] + sig0[3] + sig0[5]) * (1.f / 3.f) + (float)C->K * inc;

    const double dt = cc.dt;
    const float TwoGf = (float)C->twoG;
    const double TwoG32 = kSqrt32 * C->twoG;
    const double heatFrac = C->heatFrac;
    float s[6];
    const float pm = (sig0[0] + sig0[3] + sig0[5]) * (1.f / 3.f);
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = sig0[c];
    s[0] -= pm;
    s[3] -= pm;
    s[5] -= pm;
    const float Snorm0 = __builtin_amdgcn_sqrtf(sym_dot<B2>(s, s));
    const float em = (dE[0] + dE[3] + dE[5]) * (1.f / 3.f);
    s[0] += TwoGf * (dE[0] - em);
    s[1] += TwoGf * dE[1];
    if constexpr (!B2) s[2] += TwoGf * dE[2];
    s[3] += TwoGf * (dE[3] - em);
    if constexpr (!B2) s[4] += TwoGf * dE[4];
    s[5] += TwoGf * (dE[5] - em);
    const float Snorm = __builtin_amdgcn_sqrtf(sym_dot<B2>(s, s));
    // Work of the increment, 1/2 dE : (old + new) with new = scale s + p I
    // (the deviator scaled by the return, plus the pressure), from three
    // sums formed now; and the final rotation (Element.C:1146-1161) of the
    // trial deviator: the new stress R (scale s + p I) R^T is
    // scale R s R^T + p I for the orthogonal R.  dE, the old stress and R
    // are then not kept through the radial return (its FP64 law evaluation
    // is the register peak of the kernel).
    // power = 1/2 (work_a + scale work_s) once the return has fixed scale
    float work_a = sym_dot<B2>(dE, sig0) + pressure * (dE[0] + dE[3] + dE[5]), work_s = sym_dot<B2>(dE, s);
    float Rs[6];
#pragma unroll
    for (int c = 0; c < 6; c++) Rs[c] = s[c];
    rotate_sym_f32<B2>(R, Rs);
    // Element::computeDensity (Element.C:1243-1256)
    if (cc.density_due) {
      double rr = C->rho0 * (double)at(gbl(C->detJ0f) + g0, 4 * ip) / (double)detJ;
      if constexpr (AXI) rr /= (double)radius;
      at(gbl(C->rho) + g0, 8 * ip) = rr;
    }
    // trial test and radial return (the yield stress is stored in FP32)
    const float Strialf = (float)kSqrt32 * Snorm;
    float yieldf = sy0;
    double gamma = 0.0;
    float scale = 1.f;
    const int law = C->law;
    double lp[8];
#pragma unroll
    for (int k = 0; k < 8; k++) lp[k] = C->lp[k];
    if (yieldf == 0.f) {
      if constexpr (XL) {
        bool ok0 = true;
        yieldf = (float)(law == DNLG_LAW_ANN2SIG
                             ? law_yield_of<true>(C, law, lp, kGammaInitial, kGammaInitial / dt, Tcur, dt, ok0)
                             : law_yield_of<false>(C, law, lp, kGammaInitial, kGammaInitial / dt, Tcur, dt, ok0));
        if (!ok0) set_status(J, DNLG_E_TABLE_RANGE, par);
      } else {
        yieldf = (float)law_yield(law, lp, kGammaInitial, kGammaInitial / dt, Tcur);
      }
      at(gbl(C->sy32) + g0, 4 * ip) = yieldf;
    }
    if (Strialf > yieldf) {
      // radial return (Element.C:485-601): Newton on gamma in FP32, then
      // (RR64) the FP64 Newton resumed where it leaves off; ep, T and the
      // energies accumulate in FP64.  What the stress update needs after
      // the return waits in LDS meanwhile (the return's FP64 law evaluation
      // is the kernel's register peak); the empty asm keeps the compiler
      // from holding the values in registers across it.
      s_keep[0 * 64 + lane] = work_a;
      s_keep[1 * 64 + lane] = work_s;
      s_keep[2 * 64 + lane] = pressure;
#pragma unroll
      for (int c = 0; c < 6; c++)
        if (!B2 || (c != 2 && c != 4)) s_keep[(3 + c) * 64 + lane] = Rs[c];
      asm volatile("" ::: "memory");
      float g;
      bool ok = true;
      const double rdt = 1.0 / dt;  // one division for the return's rates
      if constexpr (XL) {
        bool in_table = true;
        auto law_at = [&](double epsp, double depsp, double T, double &y, double &h) {
          if (law == DNLG_LAW_ANN2SIG)
            law_eval<true>(C, law, lp, heatFrac, epsp, depsp, T, dt, y, h, in_table);
          else
            law_eval<false>(C, law, lp, heatFrac, epsp, depsp, T, dt, y, h, in_table);
          return in_table;
        };
        double y = (double)yieldf;  // the stored yield stress
        if (C->stress_method == DNLG_STRESS_DIRECT) {
          // Element::computeStressDirect (Element.C:625-768): one linearised
          // return from the stored yield stress and the slope at the stored
          // state (plastic strain, or gammaInitial before the first yield)
          double y0, hard;
          const double dep0 = (double)at(gbl(C->dep32) + g0, 4 * ip);
          law_at(ep0 > 0.0 ? ep0 : kGammaInitial, dep0, Tcur, y0, hard);
          const double TwoG = C->twoG;
          gamma = ((double)Snorm - kSqrt23 * y) / (TwoG * (1.0 + hard / (1.5 * TwoG)));
          const double T = Tcur + 0.5 * gamma * heatFrac * (kSqrt23 * y + (double)Snorm0);
          const double epsp = ep0 + kSqrt23 * gamma, depsp = kSqrt23 * gamma / dt;
          y = law == DNLG_LAW_ANN2SIG ? law_yield_of<true>(C, law, lp, epsp, depsp, T, dt, in_table)
                                      : law_yield_of<false>(C, law, lp, epsp, depsp, T, dt, in_table);
        } else {
          gamma = ep0 == 0.0 ? kSqrt32 * kGammaInitial : (double)gam0;
          ok = radial_return_fp64(law_at, C->nr_max_it, heatFrac, TwoG32, (double)Strialf, (double)Snorm0, Tcur, ep0,
                                  rdt, 0.0, ((double)Strialf - y) / TwoG32, y, gamma);
        }
        if (!in_table) set_status(J, DNLG_E_TABLE_RANGE, par);
        g = (float)gamma;
        yieldf = (float)y;
      } else {
        float pf[8];
#pragma unroll
        for (int k = 0; k < 8; k++) pf[k] = C->lpf[k];
        const LawF L = law_f32_setup(law, pf, C->inv_dTf, (float)dt, (float)heatFrac);
        g = gam0;
        NewtonState last;
        ok = radial_return_fp32(L, C->nr_max_it, Strialf, (float)TwoG32, (float)heatFrac, (float)Tcur, ep0, Snorm0,
                                yieldf, g, last);
        gamma = (double)g;
        if constexpr (RR64) {
          double y = (double)last.y;
          gamma = (double)last.g;
          auto law_at = [&](double epsp, double depsp, double T, double &yl, double &h) {
            law_yield_and_slope<ShortOps>(law, lp, heatFrac, epsp, depsp, T, dt, yl, h);
            return true;
          };
          ok = radial_return_fp64(law_at, C->nr_max_it, heatFrac, TwoG32, (double)Strialf, (double)Snorm0, Tcur, ep0,
                                  rdt, (double)last.gmin, (double)last.gmax, y, gamma);
          g = (float)gamma;
          yieldf = (float)y;
        }
      }
      if (!ok) set_status(J, DNLG_E_NR_DIVERGED, par);
      asm volatile("" ::: "memory");
      work_a = s_keep[0 * 64 + lane];
      work_s = s_keep[1 * 64 + lane];
      pressure = s_keep[2 * 64 + lane];
#pragma unroll
      for (int c = 0; c < 6; c++)
        if (!B2 || (c != 2 && c != 4)) Rs[c] = s_keep[(3 + c) * 64 + lane];
      scale = 1.f - TwoGf * g * __builtin_amdgcn_rcpf(Snorm);
      at(gbl(C->ep) + g0, 8 * ip) = ep0 + kSqrt23 * gamma;
      at(gbl(C->dep32) + g0, 4 * ip) = (float)(kSqrt23 * gamma * rdt);
      at(gbl(C->gam32) + g0, 4 * ip) = g;
      at(gbl(C->sy32) + g0, 4 * ip) = yieldf;
    }
    const float power = 0.5f * (work_a + scale * work_s);
    at(gbl(C->eint) + g0, 8 * ip) += (double)power * C->inv_rho0;
    if (gamma != 0.0) {
      const float sn = scale * Snorm;  // |scale s|
      const double w = 0.5 * gamma * ((double)sn + (double)Snorm0);
      at(gbl(C->einel) + g0, 8 * ip) += w * C->inv_rho0;
      at(gbl(C->T) + g0, 8 * ip) = Tcur + heatFrac * w;
    }
#pragma unroll
    for (int c = 0; c < 6; c++) sig[c] = scale * Rs[c];
    sig[0] += pressure;
    sig[3] += pressure;
    sig[5] += pressure;
#pragma unroll
    for (int c = 0; c < 6; c++)
      if (!B2 || (c != 2 && c != 4)) at(gbl(C->sig32) + c * G + g0, 4 * ip) = sig[c];
  } else {
#pragma unroll
    for (int c = 0; c < 6; c++) sig[c] = 0.f;
  }
  // ---- internal forces (Element.C:362-404): T = w J^-1 sigma per Gauss
  // point, J^-1 and det J read back from LDS (they are not kept in
  // registers through the radial return); the empty asm keeps the compiler
  // from holding them in registers instead
  asm volatile("" ::: "memory");
  float invf[DIM][DIM];
#pragma unroll
  for (int k = 0; k < DIM; k++)
#pragma unroll
    for (int j = 0; j < DIM; j++) invf[k][j] = s_jinv[(k * DIM + j) * 64 + lane];
  float w = s_jinv[DIM * DIM * 64 + lane];
  if constexpr (AXI) w *= (float)k2PI * radius;
#pragma unroll
  for (int k = 0; k < DIM; k++)
#pragma unroll
    for (int i = 0; i < DIM; i++) {
      float t = 0.f;
#pragma unroll
      for (int j = 0; j < DIM; j++) {
        const int si = (i == j) ? (i == 0 ? 0 : (i == 1 ? 3 : 5)) : (i + j == 1 ? 1 : (i + j == 2 ? 2 : 4));
        t += invf[k][j] * sig[si];
      }
      T9[k][i] = valid ? w * t : 0.f;
    }
  if constexpr (AXI) hoop = valid ? w * sig[5] * __builtin_amdgcn_rcpf(radius) : 0.f;
  // Walsh transform of T over the Gauss points, then node lane b combines
  // four (two) of them per direction, read from the element's lanes in LDS
  float WT[DIM][DIM];
#pragma unroll
  for (int k = 0; k < DIM; k++)
#pragma unroll
    for (int i = 0; i < DIM; i++) WT[k][i] = wht<NB>(T9[k][i], cs);
  if constexpr (DIM == 3) {
#pragma unroll
    for (int k = 0; k < 3; k++) s_wt[ForceRows::at(lw, b, k)] = make_float4(WT[k][0], WT[k][1], WT[k][2], 0.f);
  } else {
    s_wt[ForceRows::at(lw, b, 0)] = make_float4(WT[0][0], WT[0][1], WT[1][0], WT[1][1]);
    if constexpr (AXI) s_wt[ForceRows::at(lw, b, 1)] = make_float4(wht<NB>(hoop, cs), 0.f, 0.f, 0.f);
  }
  __builtin_amdgcn_wave_barrier();
  float f[DIM];
#pragma unroll
  for (int i = 0; i < DIM; i++) f[i] = 0.f;
#pragma unroll
  for (int k = 0; k < DIM; k++) {
    const float sk = sgn(b, k);
    if constexpr (DIM == 3) {
      const int l = (k + 1) % 3, m = (k + 2) % 3;
      const float cl = g * sgn(b, l), cm = g * sgn(b, m), clm = cl * cm;
      const float4 a0 = s_wt[ForceRows::at(lw, 0, k)];
      const float4 al = s_wt[ForceRows::at(lw, 1 << l, k)];
      const float4 am = s_wt[ForceRows::at(lw, 1 << m, k)];
      const float4 alm = s_wt[ForceRows::at(lw, (1 << l) + (1 << m), k)];
      f[0] -= sk * (a0.x + cl * al.x + cm * am.x + clm * alm.x);
      f[1] -= sk * (a0.y + cl * al.y + cm * am.y + clm * alm.y);
      f[2] -= sk * (a0.z + cl * al.z + cm * am.z + clm * alm.z);
    } else {
      const int l = 1 - k;
      const float cl = g * sgn(b, l);
      const float4 a0 = s_wt[ForceRows::at(lw, 0, 0)];
      const float4 al = s_wt[ForceRows::at(lw, 1 << l, 0)];
      if (k == 0) {
        f[0] -= sk * (a0.x + cl * al.x);
        f[1] -= sk * (a0.y + cl * al.y);
      } else {
        f[0] -= sk * (a0.z + cl * al.z);
        f[1] -= sk * (a0.w + cl * al.w);
      }
    }
  }
  if constexpr (AXI) {
    // hoop term: f_r -= sum_p N_a(p) w_p sigma_tt / r_p
    const float h0 = s_wt[ForceRows::at(lw, 0, 1)].x, h1 = s_wt[ForceRows::at(lw, 1, 1)].x;
    const float h2 = s_wt[ForceRows::at(lw, 2, 1)].x, h3 = s_wt[ForceRows::at(lw, 3, 1)].x;
    const float c0 = g * sgn(b, 0), c1 = g * sgn(b, 1);
    f[0] -= h0 + c0 * h1 + c1 * h2 + c0 * c1 * h3;
  }
  if (valid) {
    DNLG_GLOBAL float *out = gbl(C->elemF32) + (int64_t)slot * DIM;
#pragma unroll
    for (int i = 0; i < DIM; i++) out[i] = f[i] * inv2D;
  }
}

}  // namespace dnlg
