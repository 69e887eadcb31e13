// elem_mixed.h -- mixed-precision element kernel (included by kernels.hip).
//
// Same phases as k_elem_fp64 (Element.C:1103-1161, 362-404, 407-622,
// 1243-1256) with FP32 geometry, strain, stress and forces, and FP64 for the
// radial-return Newton iteration on gamma, the equivalent plastic strain, the
// temperature and the energy accumulators.
//
// Strain increment in increment form: with H = d(du)/dx the right
// Cauchy-Green tensor is C = I + X, X = H + H^T + H^T H (small), and
//   ln U = 1/2 ln(I + X)   = 1/2 sum_k (-1)^(k+1) X^k / k,
//   U^-1 = (I + X)^(-1/2)  = sum_k binom(-1/2, k) X^k,
//   R = F U^-1 = (I + H) U^-1,
// evaluated with powers of X up to X^6; for ||X||_F <= 0.05 the truncation
// is below 1e-9 relative, i.e. under FP32 rounding.  Larger increments take
// the FP64 eigen-decomposition path (polar_lnU_fp64).  This avoids the
// catastrophic cancellation of forming F^T F - I and ln(lambda ~ 1) in FP32.

namespace dnlg {

__device__ __forceinline__ void sym_mul(const float A[6], const float B[6], float C[6]) {
  // C = A B for commuting symmetric A, B (powers of one matrix): symmetric
  C[0] = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
  C[1] = A[0] * B[1] + A[1] * B[3] + A[2] * B[4];
  C[2] = A[0] * B[2] + A[1] * B[4] + A[2] * B[5];
  C[3] = A[1] * B[1] + A[3] * B[3] + A[4] * B[4];
  C[4] = A[1] * B[2] + A[3] * B[4] + A[4] * B[5];
  C[5] = A[2] * B[2] + A[4] * B[4] + A[5] * B[5];
}

// Large increments: FP64 eigen path (rarely taken).
__device__ __forceinline__ void polar_lnU_large(const float *H, float *dE, float *R) {
  double F[9], e64[6], r64[9];
#pragma unroll
  for (int k = 0; k < 9; k++) F[k] = (double)H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
  polar_lnU_fp64(F, e64, r64);
#pragma unroll
  for (int c = 0; c < 6; c++) dE[c] = (float)e64[c];
#pragma unroll
  for (int c = 0; c < 9; c++) R[c] = (float)r64[c];
}

// H (row-major 3x3, FP32) -> dE = ln U (sym), R = F U^-1 (row-major).
__device__ __forceinline__ void polar_lnU_mixed(const float H[9], float dE[6], float R[9]) {
  float X[6];
  X[0] = 2.f * H[0] + (H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
  X[1] = H[1] + H[3] + (H[0] * H[1] + H[3] * H[4] + H[6] * H[7]);
  X[2] = H[2] + H[6] + (H[0] * H[2] + H[3] * H[5] + H[6] * H[8]);
  X[3] = 2.f * H[4] + (H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
  X[4] = H[5] + H[7] + (H[1] * H[2] + H[4] * H[5] + H[7] * H[8]);
  X[5] = 2.f * H[8] + (H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
  const float nrm2 = X[0] * X[0] + X[3] * X[3] + X[5] * X[5] + 2.f * (X[1] * X[1] + X[2] * X[2] + X[4] * X[4]);
  if (nrm2 > 0.05f * 0.05f) {
    polar_lnU_large(H, dE, R);
    return;
  }
  float X2[6], X3[6], X4[6], X5[6], X6[6];
  sym_mul(X, X, X2);
  sym_mul(X2, X, X3);
  sym_mul(X2, X2, X4);
  sym_mul(X4, X, X5);
  sym_mul(X3, X3, X6);
  float Ui[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    dE[c] = 0.5f * (X[c] - X2[c] * (1.f / 2.f) + X3[c] * (1.f / 3.f) - X4[c] * (1.f / 4.f) +
                    X5[c] * (1.f / 5.f) - X6[c] * (1.f / 6.f));
    Ui[c] = -X[c] * (1.f / 2.f) + X2[c] * (3.f / 8.f) - X3[c] * (5.f / 16.f) + X4[c] * (35.f / 128.f) -
            X5[c] * (63.f / 256.f) + X6[c] * (231.f / 1024.f);
  }
  Ui[0] += 1.f;
  Ui[3] += 1.f;
  Ui[5] += 1.f;
  // R = (I + H) Ui
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

template <int NN>
__device__ __forceinline__ float tabDf(int p, int a, int j) {
  if constexpr (NN == 8) return c_hexDf[p][a][j];
  else return c_quaDf[p][a][j];
}
template <int NN>
__device__ __forceinline__ float tabNf(int p, int a) {
  if constexpr (NN == 8) return c_hexNf[p][a];
  else return c_quaNf[p][a];
}

template <int NN, int DIM, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_elem_mixed(JobDev *const *jobs, const Task *tasks) {
  constexpr int NIP = NN;
  constexpr int EPB = BLOCK / NIP;
  __shared__ float s_x[EPB][NN][DIM];  // coordinates relative to node 0
  __shared__ float s_u[EPB][NN][DIM];
  __shared__ float s_dN[EPB][NIP][NN][DIM];
  __shared__ float s_sw[EPB][NIP][8];
  __shared__ float s_tr[EPB][NIP];

  const Task tk = tasks[blockIdx.x];
  JobDev *J = jobs[tk.job];
  if (!J->c.running || J->c.status != 0) return;
  const int le = threadIdx.x / NIP;
  const int p = threadIdx.x % NIP;
  const int64_t e = tk.first + le;
  const bool valid = e < tk.end;
  const int64_t N = J->N;
  const int f0 = J->c.f0;
  double x0r = 0.0;  // radius of node 0 (axisymmetric)
  if (valid) {
    const int nd = J->conn[e * NN + p];
    const int n0 = J->conn[e * NN];
    const double *u0 = J->fu[f0];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      s_x[le][p][d] = (float)(J->x[d * N + nd] - J->x[d * N + n0]);
      s_u[le][p][d] = (float)u0[d * N + nd];
    }
    if constexpr (DIM == 2) x0r = J->x[n0];
  }
  __syncthreads();

  const int64_t G = J->G;
  const int64_t g = e * NIP + p;
  float dN[NN][DIM];
  float detJ = 1.f, radius = 1.f;
  float dE[6] = {0, 0, 0, 0, 0, 0}, R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (valid) {
    float Jm[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int j = 0; j < DIM; j++) Jm[i][j] = 0.f;
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) Jm[i][j] += tabDf<NN>(p, a, j) * s_x[le][a][i];
    float inv[DIM][DIM];
    if constexpr (DIM == 3) {
      const float *dd = &Jm[0][0];
      detJ = dd[1] * dd[5] * dd[6] - dd[2] * dd[4] * dd[6] + dd[2] * dd[3] * dd[7] - dd[0] * dd[5] * dd[7] -
             dd[1] * dd[3] * dd[8] + dd[0] * dd[4] * dd[8];
      const float id = 1.f / detJ;
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
      const float id = 1.f / detJ;
      inv[0][0] = Jm[1][1] * id;
      inv[0][1] = -Jm[0][1] * id;
      inv[1][0] = -Jm[1][0] * id;
      inv[1][1] = Jm[0][0] * id;
    }
    if (detJ < 0.f) set_status(J, DNLG_E_NEG_JACOBIAN);
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int j = 0; j < DIM; j++) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < DIM; k++) s += tabDf<NN>(p, a, k) * inv[k][j];
        dN[a][j] = s;
      }
    float H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) H[3 * i + j] += dN[a][j] * s_u[le][a][i];
    if constexpr (DIM == 2) {
      float rl = 0.f, vr = 0.f;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        rl += tabNf<NN>(p, a) * s_x[le][a][0];
        vr += tabNf<NN>(p, a) * s_u[le][a][0];
      }
      radius = (float)(x0r + (double)rl);
      H[8] = vr / radius;
    }
    polar_lnU_mixed(H, dE, R);
    s_tr[le][p] = dE[0] + dE[3] + dE[5];
  }
  __syncthreads();

  float sig[6], wdj = 0.f;
  if (valid) {
    float inc = 0.f;
#pragma unroll
    for (int q = 0; q < NIP; q++) inc += s_tr[le][q];
    inc *= (1.f / NIP);
#pragma unroll
    for (int c = 0; c < 6; c++) sig[c] = J->sig32[c * G + g];
    const float pressure = (sig[0] + sig[3] + sig[5]) * (1.f / 3.f) + (float)J->K * inc;

    const double dt = J->c.dt;
    const float TwoGf = (float)J->twoG;
    const double TwoG32 = kSqrt32 * J->twoG;
    const double heatFrac = J->heatFrac;
    float old[6], s[6];
#pragma unroll
    for (int c = 0; c < 6; c++) old[c] = sig[c];
    const float pm = (old[0] + old[3] + old[5]) * (1.f / 3.f);
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = old[c];
    s[0] -= pm;
    s[3] -= pm;
    s[5] -= pm;
    const float Snorm0 = sqrtf(s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + 2.f * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4]));
    const float em = (dE[0] + dE[3] + dE[5]) * (1.f / 3.f);
    float de[6];
#pragma unroll
    for (int c = 0; c < 6; c++) de[c] = dE[c];
    de[0] -= em;
    de[3] -= em;
    de[5] -= em;
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] += TwoGf * de[c];
    const float Snorm = sqrtf(s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + 2.f * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4]));
    const double Strial = kSqrt32 * (double)Snorm;
    const double Tcur = J->T[g];
    double yield = (double)J->sy32[g];
    double gamma = 0.0;
    const int law = J->law;
    const double *lp = J->lp;
    if (yield == 0.0) {
      yield = law_yield(law, lp, kGammaInitial, kGammaInitial / dt, Tcur);
      J->sy32[g] = (float)yield;
    }
    if (Strial > yield) {
      // radial return in FP64 (Element.C:485-601)
      const double ep0 = J->ep[g];
      double gmin = 0.0, gmax = (Strial - yield) / TwoG32;
      gamma = (double)J->gam32[g];
      if (ep0 == 0.0) gamma = kSqrt32 * kGammaInitial;
      double depsp = kSqrt23 * gamma / dt;
      double epsp = ep0 + kSqrt23 * gamma;
      double T = Tcur + 0.5 * gamma * heatFrac * (kSqrt23 * yield + (double)Snorm0);
      int iterate = 0;
      while (true) {
        double hard;
        law_yield_and_slope(law, lp, heatFrac, epsp, depsp, T, dt, yield, hard);
        const double fun = Strial - gamma * TwoG32 - yield;
        if (fun < 0.0)
          gmax = gamma;
        else
          gmin = gamma;
        const double dfun = TwoG32 + kSqrt23 * hard;
        double dgamma = fun / dfun;
        gamma += dgamma;
        if ((gmax - gamma) * (gamma - gmin) < 0.0) {
          dgamma = 0.5 * (gmax - gmin);
          gamma = gmin + dgamma;
        }
        if (fabs(dgamma) < kTolNR) break;
        depsp = kSqrt23 * gamma / dt;
        epsp = ep0 + kSqrt23 * gamma;
        T = Tcur + 0.5 * gamma * heatFrac * (kSqrt23 * yield + (double)Snorm0);
        if (++iterate > kItMax) {
          set_status(J, DNLG_E_NR_DIVERGED);
          break;
        }
      }
      const float scale = (float)(1.0 - J->twoG * gamma / (double)Snorm);
#pragma unroll
      for (int c = 0; c < 6; c++) s[c] *= scale;
      J->ep[g] = ep0 + kSqrt23 * gamma;
      J->dep32[g] = (float)(kSqrt23 * gamma / dt);
      J->gam32[g] = (float)gamma;
      J->gamc32[g] += (float)gamma;
      J->sy32[g] = (float)yield;
    }
#pragma unroll
    for (int c = 0; c < 6; c++) sig[c] = s[c];
    sig[0] += pressure;
    sig[3] += pressure;
    sig[5] += pressure;
    const float power = 0.5f * (dE[0] * (old[0] + sig[0]) + dE[3] * (old[3] + sig[3]) + dE[5] * (old[5] + sig[5]) +
                                2.f * (dE[1] * (old[1] + sig[1]) + dE[2] * (old[2] + sig[2]) +
                                       dE[4] * (old[4] + sig[4])));
    J->eint[g] += (double)power / J->rho0;
    if (gamma != 0.0) {
      const float sn = sqrtf(s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + 2.f * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4]));
      const double w = 0.5 * gamma * ((double)sn + (double)Snorm0);
      J->einel[g] += w / J->rho0;
      J->T[g] = Tcur + heatFrac * w;
    }
    rotate_sym(R, sig);
#pragma unroll
    for (int c = 0; c < 6; c++) J->sig32[c * G + g] = sig[c];
    if (J->c.density_due) {
      double rr = J->rho0 * (double)J->detJ0f[g] / (double)detJ;
      if constexpr (DIM == 2) rr /= (double)radius;
      J->rho[g] = rr;
    }
    wdj = detJ;
    if constexpr (DIM == 2) wdj *= (float)k2PI * radius;
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int j = 0; j < DIM; j++) s_dN[le][p][a][j] = dN[a][j];
#pragma unroll
    for (int c = 0; c < 6; c++) s_sw[le][p][c] = sig[c];
    s_sw[le][p][6] = wdj;
    s_sw[le][p][7] = radius;
  }
  __syncthreads();

  if (valid) {
    const int a = p;
    float f[DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++) f[i] = 0.f;
#pragma unroll
    for (int q = 0; q < NIP; q++) {
      const float w = s_sw[le][q][6];
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) f[i] -= s_dN[le][q][a][j] * s_sw[le][q][sym_idx(j, i)] * w;
      if constexpr (DIM == 2) f[0] -= tabNf<NN>(q, a) * s_sw[le][q][5] / s_sw[le][q][7] * w;
    }
    float *out = J->elemF32 + (e * NN + a) * DIM;
#pragma unroll
    for (int i = 0; i < DIM; i++) out[i] = f[i];
  }
}

}  // namespace dnlg
