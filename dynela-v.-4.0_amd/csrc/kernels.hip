// kernels.hip -- CDNA4 kernels of one explicit cycle (FP64 element path,
// nodal kernels, Courant minimum, control) and of the pre-loop set-up.
//
// Cycle = Explicit::solve loop body (Explicit.C:184-252), up to four launches:
//   k_cycle_step    endStep + computeTimeStep of the previous cycle
//                   (Explicit.C:539-552, Solver.C:196-236), updateTimes
//                   (Explicit.C:414-427), field0 BCs of computePredictions;
//                   inside a captured sequence of cycles that share their
//                   time-step phase the corrector or the Courant kernel
//                   does this at its end instead
//   k_elem_*        computeStrains, computePressure, computeStress,
//                   computeFinalRotation, computeInternalForces,
//                   computeDensity (Model.C:800-899, Explicit.C:555-562)
//   k_correct       computePredictions + explicitSolve + BCs
//                   (Explicit.C:438-536), gather-based force assembly in
//                   element order (Model.C:864-899)
//   k_courant       computeCourantTimeStep (Model.C:636-670), per-block min
//                   (every dt_freq-th cycle; an early exit otherwise, or
//                   not launched when no running job is due)
// The element Jacobian of Explicit.C:245-246 is recomputed inside the next
// cycle's element kernel from the same coordinates, so nothing per-IP
// geometric is stored between cycles.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "material.h"
#include "polar_fp64.h"
#include "control.h"

namespace dnlg {

// Element shape tables at the Gauss points: the closed forms tabulated in
// ElQua4NAx.C:63-110 and ElHex8N3D.C:93-231, generated on the host
// (api.hip, make_tables) and uploaded once.
__constant__ double c_hexN[8][8];
__constant__ double c_hexD[8][8][3];
__constant__ double c_quaN[4][4];
__constant__ double c_quaD[4][4][2];
__constant__ float c_hexNf[8][8];
__constant__ float c_hexDf[8][8][3];
__constant__ float c_quaNf[4][4];
__constant__ float c_quaDf[4][4][2];

template <int NN>
__device__ __forceinline__ double tabN(int p, int a) {
  if constexpr (NN == 8) return c_hexN[p][a];
  else return c_quaN[p][a];
}
template <int NN>
__device__ __forceinline__ double tabD(int p, int a, int j) {
  if constexpr (NN == 8) return c_hexD[p][a][j];
  else return c_quaD[p][a][j];
}

__device__ __forceinline__ int sym_idx(int i, int j) { return (i == j) ? (i == 0 ? 0 : (i == 1 ? 3 : 5)) : (i + j == 1 ? 1 : (i + j == 2 ? 2 : 4)); }

// Sticky job status of the cycle with increment parity `par`.
__device__ __forceinline__ void set_status(JobDev *J, int st, int par) { atomicCAS(&J->sync.st[par], 0, st); }

// ---------------------------------------------------------------------------
// Cycle control.  Between two cycles a job's control record goes through
//   close  Explicit::endStep (Explicit.C:539-552): t += dt, field1 becomes
//          field0; then Solver::computeTimeStep (Solver.C:196-236, Courant
//          branch: omega_max = 2 / dt_crit) when the time step was due;
//   open   Explicit::updateTimes (Explicit.C:414-427): last-step clip,
//          increment, and the engine's per-cycle flags.
// k_cycle_step does it where a launch sequence starts, after time-step
// cycles, and (close only) where a sequence ends; between the other cycles of
// a sequence the next cycle's element kernel derives the new record itself
// (cycle_control) and the corrector has already written the new cycle's
// field0 BC values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_control(JobDev *J, int slot, const Control &c) {
  DNLG_GLOBAL Control &d = *gbl(&J->c[slot]);
  d.t = c.t;
  d.dt = c.dt;
  d.up_to = c.up_to;
  d.inc = c.inc;
  d.f0 = c.f0;
  d.running = c.running;
  d.courant_due = c.courant_due;
  d.density_due = c.density_due;
  d.preloop = c.preloop;
  d.pending = c.pending;
}

// Close the cycle that ran last (if any) and, with `open`, open the next.
// status: the job's status after that cycle; freq > 0: the maximum
// frequency of the time-step update (Courant: 2 / min L/c; power iteration:
// sqrt(fmax)), dt = safety Omega_s / freq.  Returns whether the field0 BC
// values of the new cycle differ from what the corrector left in field0
// (first cycle of an advance, a ramped speed BC, or a speed BC under a
// changed time step).
__device__ __forceinline__ bool close_open(CJobDev *C, Control &c, int status, double freq, bool open) {
  const double dt_last = c.dt;
  const bool resumed = !c.pending;
  if (c.pending && status == 0) {
    c.t += c.dt;
    c.f0 = 1 - c.f0;
    if (freq > 0.0) c.dt = C->safety * C->omegaS / freq;
  }
  c.pending = 0;
  bool bc_work = false;
  if (open) {
    c.preloop = 0;
    // a job whose power iteration stalled waits for the host-driven
    // iteration (k_pi_finish clears the flag), a held one for its result
    // file
    c.running = (status == 0 && c.t < c.up_to && !C->pi.stall && !C->res.hold) ? 1 : 0;
    if (c.running) {
      if (c.t + c.dt > c.up_to) c.dt = c.up_to - c.t;
      c.inc++;
      c.courant_due = (c.inc % C->dt_freq == 0) ? 1 : 0;
      // densities are consumed by the Courant step, read back at the end
      // and, when a history item reads them, after every cycle
      c.density_due = (C->density_always || c.courant_due || !(c.t + c.dt < c.up_to)) ? 1 : 0;
      c.pending = 1;
      bc_work = resumed || (C->bc_flags & 2) || ((C->bc_flags & 1) && c.dt != dt_last);
    }
  } else {
    c.running = 0;
  }
  return bc_work;
}

// Whether the Courant kernel computed partial minima in the cycle that ran
// last (Explicit.C:243-251: every dt_freq increments, while t_{n+1} < t_up_to).
__device__ __forceinline__ bool courant_ran(const Control &c, int status) {
  return c.pending && status == 0 && c.courant_due && c.t + c.dt < c.up_to;
}

// The record of the cycle that a launch works on.  derive == 0: slot cs
// holds it.  derive != 0 (the first launch of a cycle that no k_cycle_step
// opened, after a cycle without a time-step update): slot cs holds the
// previous cycle; every block derives the new record from it and the job's
// block 0 stores it in slot 1 - cs for the cycle's later launches.  The
// previous cycle's failures are in its own parity word, which no launch of
// the new cycle writes; a job that did not run in it stays stopped (it
// finished or failed), so reading both words is then race-free too.
// The record is uniform over the block; it is returned in scalar registers
// (the FP64 arithmetic of the derivation runs on the vector unit).
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform_l(int64_t v) {
  const unsigned lo = (unsigned)uniform_i((int)(v & 0xffffffffLL));
  const unsigned hi = (unsigned)uniform_i((int)(v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uniform_d(double v) { return __longlong_as_double(uniform_l(__double_as_longlong(v))); }

__device__ __forceinline__ Control cycle_control(JobDev *J, CJobDev *C, int cs, int derive, int blk) {
  Control c = load_control(C, cs);
  if (derive) {
    const int st0 = C->sync.st[0], st1 = C->sync.st[1];
    const int status = c.pending ? ((c.inc & 1) ? st1 : st0) : (st0 != 0 ? st0 : st1);
    close_open(C, c, status, 0.0, true);
    if (blk == 0 && threadIdx.x == 0) store_control(J, 1 - cs, c);
    c.t = uniform_d(c.t);
    c.dt = uniform_d(c.dt);
    c.inc = uniform_l(c.inc);
    c.f0 = uniform_i(c.f0);
    c.running = uniform_i(c.running);
    c.courant_due = uniform_i(c.courant_due);
    c.density_due = uniform_i(c.density_due);
    c.preloop = uniform_i(c.preloop);
    c.pending = uniform_i(c.pending);
  }
  return c;
}


// The hardening law's yield stress (and slope) at (epsp, depsp, T) for the
// FP64 radial returns: a tabular law through its table (out of it: `ok`
// cleared), the closed-form laws through material.h; ANN: the kernel
// variant for network laws.
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

// ---------------------------------------------------------------------------
// Element kernel, FP64.  One lane per integration point (NIP lanes per
// element, NIP == NN for both element types); the element's nodes are staged
// in LDS by its own lanes, per-IP inverse Jacobians, stresses and weights go back to
// LDS so that lane a can sum node a's force over the IPs in the reference's
// order (IP, then direction j).
// ---------------------------------------------------------------------------
// Three waves per SIMD (168 VGPRs, a short spill); the LDS of a 16-hex block
// (24 KB) allows three.  Measured 11 % faster than two waves without spill.
// ANN: the variant for jobs whose law is DNLG_LAW_ANN2SIG (its network
// evaluation keeps per-lane arrays that the other laws' variants do without).
template <int NN, int DIM, int BLOCK, bool AXI, bool ANN>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void k_elem_fp64(
    JobDev *const *jobs, const Task *tasks, int cs, int derive) {
  constexpr int NIP = NN;
  constexpr int EPB = BLOCK / NIP;
  __shared__ double s_x[EPB][NN][DIM];
  __shared__ double s_u[EPB][NN][DIM];
  __shared__ double s_inv[EPB][NIP][DIM][DIM];  // J^-1 per IP, for the force phase
  __shared__ double s_sw[EPB][NIP][8];
  __shared__ double s_tr[EPB][NIP];

  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  const Control cc = cycle_control(J, C, cs, derive, tk.blk);
  if (!cc.running) return;
  const int par = (int)(cc.inc & 1);
  const int le = threadIdx.x / NIP;
  const int p = threadIdx.x % NIP;
  const int64_t e = tk.first + le;
  const bool valid = e < tk.end;
  const int64_t N = C->N;
  const int f0 = cc.f0;

  // stage the element's nodes: lane p brings node p
  if (valid) {
    const int nd = gbl(C->conn)[e * NN + p];
    const DNLG_GLOBAL double *u0 = gbl(C->fu[f0]);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      s_x[le][p][d] = gbl(C->x)[d * N + nd];
      s_u[le][p][d] = u0[d * N + nd];
    }
  }
  __syncthreads();

  const int64_t G = C->G;
  const int64_t g = e * NIP + p;
  double dN[NN][DIM];
  double detJ = 1.0, radius = 1.0;
  double dE[6] = {0, 0, 0, 0, 0, 0}, R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (valid) {
    // Jacobian, inverse, B-matrix (Element3D.C:64-115, ElementPlane.C:46-102)
    double Jm[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int j = 0; j < DIM; j++) Jm[i][j] = 0.0;
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) Jm[i][j] += tabD<NN>(p, a, j) * s_x[le][a][i];
    double inv[DIM][DIM];
    if constexpr (DIM == 3) {
      const double *dd = &Jm[0][0];
      detJ = dd[1] * dd[5] * dd[6] - dd[2] * dd[4] * dd[6] + dd[2] * dd[3] * dd[7] - dd[0] * dd[5] * dd[7] -
             dd[1] * dd[3] * dd[8] + dd[0] * dd[4] * dd[8];
      const double id = 1.0 / detJ;
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
      const double id = 1.0 / detJ;
      inv[0][0] = Jm[1][1] * id;
      inv[0][1] = -Jm[0][1] * id;
      inv[1][0] = -Jm[1][0] * id;
      inv[1][1] = Jm[0][0] * id;
    }
    if (detJ < 0.0) set_status(J, DNLG_E_NEG_JACOBIAN, par);
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int j = 0; j < DIM; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; k++) s += tabD<NN>(p, a, k) * inv[k][j];
        dN[a][j] = s;
      }
#pragma unroll
    for (int k = 0; k < DIM; k++)
#pragma unroll
      for (int j = 0; j < DIM; j++) s_inv[le][p][k][j] = inv[k][j];
    // deformation gradient from the field0 increment (Element3D.C:172-195,
    // ElementAxi.C:78-100) and polar / log strain (Tensor2.C:1442-1457)
    double F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) F[3 * i + j] += dN[a][j] * s_u[le][a][i];
    if constexpr (AXI) {
      double r = 0.0, vr = 0.0;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        r += tabN<NN>(p, a) * s_x[le][a][0];
        vr += tabN<NN>(p, a) * s_u[le][a][0];
      }
      radius = r;
      F[8] += vr / radius;
    }
    polar_lnU_fp64(F, dE, R);
    s_tr[le][p] = dE[0] + dE[3] + dE[5];
  }
  __syncthreads();

  double sig[6], wdj = 0.0;
  if (valid) {
    // Element::computePressure (Element.C:1125-1143)
    double inc = 0.0;
#pragma unroll
    for (int q = 0; q < NIP; q++) inc += s_tr[le][q];
    inc /= NIP;
#pragma unroll
    for (int c = 0; c < 6; c++) sig[c] = gbl(C->sig)[c * G + g];
    const double pressure = (sig[0] + sig[3] + sig[5]) / 3.0 + C->K * inc;

    // Element::computeStress (Element.C:407-622)
    const double dt = cc.dt;
    const double TwoG = C->twoG;
    const double TwoG32 = kSqrt32 * TwoG;
    const double heatFrac = C->heatFrac;
    double old[6], s[6];
#pragma unroll
    for (int c = 0; c < 6; c++) old[c] = sig[c];
    const double pm = (old[0] + old[3] + old[5]) / 3.0;
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = old[c];
    s[0] -= pm;
    s[3] -= pm;
    s[5] -= pm;
    const double Snorm0 = sqrt(sqd(s[0]) + sqd(s[3]) + sqd(s[5]) + 2.0 * (sqd(s[1]) + sqd(s[2]) + sqd(s[4])));
    const double em = (dE[0] + dE[3] + dE[5]) / 3.0;
    double de[6];
#pragma unroll
    for (int c = 0; c < 6; c++) de[c] = dE[c];
    de[0] -= em;
    de[3] -= em;
    de[5] -= em;
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] += TwoG * de[c];
    const double Snorm = sqrt(sqd(s[0]) + sqd(s[3]) + sqd(s[5]) + 2.0 * (sqd(s[1]) + sqd(s[2]) + sqd(s[4])));
    const double Strial = kSqrt32 * Snorm;
    const double Tcur = gbl(C->T)[g];
    const double ep0 = gbl(C->ep)[g];
    double yield = gbl(C->sy)[g];
    double gamma = 0.0;
    const int law = C->law;
    double lp[8];
#pragma unroll
    for (int k = 0; k < 8; k++) lp[k] = C->lp[k];
    bool in_table = true;
    if (yield == 0.0) {
      yield = law_yield_of<ANN>(C, law, lp, kGammaInitial, kGammaInitial / dt, Tcur, dt, in_table);
      gbl(C->sy)[g] = yield;
    }
    const bool plastic = Strial > yield;
    if (plastic && C->stress_method == DNLG_STRESS_DIRECT) {
      // Element::computeStressDirect (Element.C:625-768): one linearised
      // return from the stored yield stress and the slope at the stored
      // state (plastic strain, or gammaInitial before the first yield)
      double y0, hard;
      law_eval<ANN>(C, law, lp, heatFrac, ep0 > 0.0 ? ep0 : kGammaInitial, gbl(C->dep)[g], Tcur, dt, y0, hard,
                    in_table);
      gamma = (Snorm - kSqrt23 * yield) / (TwoG * (1.0 + hard / (1.5 * TwoG)));
      const double T = Tcur + 0.5 * gamma * heatFrac * (kSqrt23 * yield + Snorm0);
      yield = law_yield_of<ANN>(C, law, lp, ep0 + kSqrt23 * gamma, kSqrt23 * gamma / dt, T, dt, in_table);
    } else if (plastic) {
      double gmin = 0.0, gmax = (Strial - yield) / TwoG32;
      gamma = gbl(C->gam)[g];
      if (ep0 == 0.0) gamma = kSqrt32 * kGammaInitial;
      double depsp = kSqrt23 * gamma / dt;
      double epsp = ep0 + kSqrt23 * gamma;
      double T = Tcur + 0.5 * gamma * heatFrac * (kSqrt23 * yield + Snorm0);
      const double eta_rcp = heatFrac;
      int iterate = 0;
      while (true) {
        double hard;
        law_eval<ANN>(C, law, lp, eta_rcp, epsp, depsp, T, dt, yield, hard, in_table);
        if (!in_table) break;
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
        T = Tcur + 0.5 * gamma * heatFrac * (kSqrt23 * yield + Snorm0);
        if (++iterate > C->nr_max_it) {
          set_status(J, DNLG_E_NR_DIVERGED, par);
          break;
        }
      }
    }
    if (!in_table) set_status(J, DNLG_E_TABLE_RANGE, par);
    if (plastic) {
      if (C->track) {
#pragma unroll
        for (int c = 0; c < 6; c++) gbl(C->epsP)[c * G + g] += gamma * s[c] / Snorm;
      }
      const double scale = 1.0 - TwoG * gamma / Snorm;
#pragma unroll
      for (int c = 0; c < 6; c++) s[c] *= scale;
      gbl(C->ep)[g] = ep0 + kSqrt23 * gamma;
      gbl(C->dep)[g] = kSqrt23 * gamma / dt;
      gbl(C->gam)[g] = gamma;
      gbl(C->gamc)[g] += gamma;
      gbl(C->sy)[g] = yield;
    }
#pragma unroll
    for (int c = 0; c < 6; c++) sig[c] = s[c];
    sig[0] += pressure;
    sig[3] += pressure;
    sig[5] += pressure;
    // specific internal energy, plastic work, adiabatic heating
    const double power = 0.5 * (dE[0] * (old[0] + sig[0]) + dE[3] * (old[3] + sig[3]) + dE[5] * (old[5] + sig[5]) +
                                2.0 * (dE[1] * (old[1] + sig[1]) + dE[2] * (old[2] + sig[2]) +
                                       dE[4] * (old[4] + sig[4])));
    gbl(C->eint)[g] += power / C->rho0;
    if (gamma != 0.0) {
      const double sn = sqrt(sqd(s[0]) + sqd(s[3]) + sqd(s[5]) + 2.0 * (sqd(s[1]) + sqd(s[2]) + sqd(s[4])));
      const double w = 0.5 * gamma * (sn + Snorm0);
      gbl(C->einel)[g] += w / C->rho0;
      gbl(C->T)[g] = Tcur + heatFrac * w;
    }
    // Element::computeFinalRotation (Element.C:1146-1161)
    rotate_sym(R, sig);
#pragma unroll
    for (int c = 0; c < 6; c++) gbl(C->sig)[c * G + g] = sig[c];
    if (C->track) {
      double t6[6];
#pragma unroll
      for (int c = 0; c < 6; c++) t6[c] = gbl(C->eps)[c * G + g] + dE[c];
      rotate_sym(R, t6);
#pragma unroll
      for (int c = 0; c < 6; c++) gbl(C->eps)[c * G + g] = t6[c];
#pragma unroll
      for (int c = 0; c < 6; c++) t6[c] = gbl(C->epsP)[c * G + g];
      rotate_sym(R, t6);
#pragma unroll
      for (int c = 0; c < 6; c++) gbl(C->epsP)[c * G + g] = t6[c];
    }
    // Element::computeDensity (Element.C:1243-1256), on x_n
    if (cc.density_due) {
      double rr = C->rho0 * gbl(C->detJ0)[g] / detJ;
      if constexpr (AXI) rr /= radius;
      gbl(C->rho)[g] = rr;
    }
    wdj = detJ;  // Gauss weights are 1 (ElQua4NAx.C:63-110, ElHex8N3D.C:93-231)
    if constexpr (AXI) wdj *= k2PI * radius;
#pragma unroll
    for (int c = 0; c < 6; c++) s_sw[le][p][c] = sig[c];
    s_sw[le][p][6] = wdj;
    s_sw[le][p][7] = radius;
  }
  __syncthreads();

  // Element::computeInternalForces (Element.C:362-404): lane a sums node a
  if (valid) {
    const int a = p;
    double f[DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++) f[i] = 0.0;
#pragma unroll
    for (int q = 0; q < NIP; q++) {
      const double w = s_sw[le][q][6];
      // node a's shape derivatives at IP q, formed as in the strain phase
      double dNq[DIM];
#pragma unroll
      for (int j = 0; j < DIM; j++) {
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; k++) d += tabD<NN>(q, a, k) * s_inv[le][q][k][j];
        dNq[j] = d;
      }
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) f[i] -= dNq[j] * s_sw[le][q][sym_idx(j, i)] * w;
      if constexpr (AXI) f[0] -= tabN<NN>(q, a) * s_sw[le][q][5] / s_sw[le][q][7] * w;
    }
    DNLG_GLOBAL double *out = gbl(C->elemF) + (int64_t)gbl(C->e2slot)[e * NN + a] * DIM;
#pragma unroll
    for (int i = 0; i < DIM; i++) out[i] = f[i];
  }
}

}  // namespace dnlg

#include "elem_mixed.h"

namespace dnlg {

// ---------------------------------------------------------------------------
// Nodal kernels
// ---------------------------------------------------------------------------
// BoundaryRestrain / BoundarySpeed ::applyConstantOnNewFields
// (BoundaryRestrain.C:98-150, BoundarySpeed.C:139-226) on both fields.
template <int DIM>
__device__ __forceinline__ void apply_bcs(CJobDev *C, int64_t n, double time, double h, double ua[DIM],
                                          double va[DIM], double aa[DIM], double ub[DIM], double vb[DIM],
                                          double ab[DIM]) {
  const int k0 = gbl(C->bc_ptr)[n], k1 = gbl(C->bc_ptr)[n + 1];
  for (int k = k0; k < k1; k++) {
    const DNLG_GLOBAL Boundary &B = gbl(C->bnd)[gbl(C->bc_idx)[k]];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      if (B.kind == DNLG_BC_RESTRAIN) {
        if (B.v[d] > 0) {
          ua[d] = ub[d] = 0.0;
          va[d] = vb[d] = 0.0;
          aa[d] = ab[d] = 0.0;
        }
      } else if (B.v[d] != 0.0) {
        if (B.ramp == DNLG_RAMP_NONE) {
          aa[d] = ab[d] = 0.0;
          va[d] = vb[d] = B.v[d];
          ua[d] = va[d] * h;
          ub[d] = vb[d] * h;
        } else {
          const double r1 = ramp_value(B.ramp, B.rmin, B.rmax, time + h);
          const double r0 = ramp_value(B.ramp, B.rmin, B.rmax, time);
          aa[d] = (r1 - r0) * B.v[d] / h;
          ab[d] = aa[d];
          va[d] = r0 * B.v[d];
          vb[d] = r1 * B.v[d];
          ua[d] = va[d] * h;
          ub[d] = vb[d] * h;
        }
      }
    }
  }
}

// Explicit::computePredictions + Explicit::explicitSolve (Explicit.C:438-536)
// in one pass per node: the predicted field1 values are formed from field0
// in registers (they are read nowhere else), the element forces of the node
// are read from its contiguous incidence slots in element order (same
// summation order as the reference's scatter-add of Model.C:864-899), a = f/M,
// Chung-Hulbert correction, BCs, coordinates (the node displacement of
// Explicit.C:530-531 is x - x0, formed at read-back).
// fold != 0 (a cycle without a time-step update, followed by the next
// cycle's element pass with derive != 0): a BC node's thread also writes the
// next cycle's field0 BC values, which k_cycle_step would otherwise write.
// Adds the forces of incidence slots k .. k+B-1 to f in slot order, with the
// B slots' loads issued together.
template <int B, int DIM, typename FT>
__device__ __forceinline__ void add_slots(const DNLG_GLOBAL FT *ef, int64_t k, double f[DIM]) {
  FT v[B][DIM];
#pragma unroll
  for (int i = 0; i < B; i++)
#pragma unroll
    for (int d = 0; d < DIM; d++) v[i][d] = ef[(k + i) * DIM + d];
#pragma unroll
  for (int i = 0; i < B; i++)
#pragma unroll
    for (int d = 0; d < DIM; d++) f[d] += (double)v[i][d];
}

template <int DIM, typename FT>
__global__ __launch_bounds__(kNodeBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_correct(
    JobDev *const *jobs, const Task *tasks, int cs, int fold) {
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  const Control c = load_control(C, cs);
  const int st0 = C->sync.st[0], st1 = C->sync.st[1];
  if (!c.running || ((c.inc & 1) ? st1 : st0) != 0) return;
  const int64_t n = tk.first + threadIdx.x;
  if (n >= tk.end) return;
  const int64_t N = C->N;
  const int a = c.f0, b = 1 - a;
  const double h = c.dt;
  const DNLG_GLOBAL FT *ef;
  if constexpr (sizeof(FT) == 8) ef = gbl(C->elemF); else ef = gbl(C->elemF32);
  const int k0 = gbl(C->n2e_ptr)[n], k1 = gbl(C->n2e_ptr)[n + 1];
  // the node's own values first, then the incidence slots in batches
  const double m = gbl(C->mass)[n];
  double va[DIM], aa[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    va[d] = gbl(C->fv[a])[d * N + n];
    aa[d] = gbl(C->fa[a])[d * N + n];
  }
  double f[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) f[d] = 0.0;
  int64_t k = k0;
  if constexpr (sizeof(FT) == 4) {
    for (; k + 8 <= k1; k += 8) add_slots<8, DIM>(ef, k, f);
    if (k + 4 <= k1) {
      add_slots<4, DIM>(ef, k, f);
      k += 4;
    }
  } else {  // FP64 slots: batches of 4 keep the kernel within its registers
    for (; k + 4 <= k1; k += 4) add_slots<4, DIM>(ef, k, f);
  }
  if (k + 2 <= k1) {
    add_slots<2, DIM>(ef, k, f);
    k += 2;
  }
  if (k < k1) add_slots<1, DIM>(ef, k, f);
  const double alphaM = C->alphaM, gam = C->gammaCH, beta = C->beta;
  double ua[DIM], ub[DIM], vb[DIM], ab[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    // predictor (Explicit.C:458-466)
    ub[d] = h * (va[d] + (0.5 - beta) * h * aa[d]);
    vb[d] = va[d] + (1.0 - gam) * h * aa[d];
    // corrector (Explicit.C:508-521)
    ab[d] = f[d] / m;
    ab[d] -= alphaM * aa[d];
    ab[d] /= (1.0 - alphaM);
    vb[d] += gam * h * ab[d];
    ub[d] += beta * (h * h) * ab[d];
  }
  const bool has_bc = gbl(C->bc_ptr)[n + 1] > gbl(C->bc_ptr)[n];
  if (has_bc) {
#pragma unroll
    for (int d = 0; d < DIM; d++) ua[d] = gbl(C->fu[a])[d * N + n];
    apply_bcs<DIM>(C, n, c.t, h, ua, va, aa, ub, vb, ab);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      gbl(C->fu[a])[d * N + n] = ua[d];
      gbl(C->fv[a])[d * N + n] = va[d];
      gbl(C->fa[a])[d * N + n] = aa[d];
    }
  }
  bool finite = true;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const double xn = gbl(C->x)[d * N + n] + ub[d];
    gbl(C->x)[d * N + n] = xn;
    finite = finite && isfinite(xn) && isfinite(vb[d]);
  }
  if (has_bc && fold) {
    // the next cycle's field0 BCs on field b, as k_cycle_step would write
    // them if the job keeps running (if it stops, field b is not read)
    Control nx = c;
    if (close_open(C, nx, 0, 0.0, true)) {
      double u2[DIM], v2[DIM], a2[DIM];
      apply_bcs<DIM>(C, n, nx.t, nx.dt, ub, vb, ab, u2, v2, a2);
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    gbl(C->fu[b])[d * N + n] = ub[d];
    gbl(C->fv[b])[d * N + n] = vb[d];
    gbl(C->fa[b])[d * N + n] = ab[d];
  }
  if (!finite) set_status(J, DNLG_E_NONFINITE, (int)(c.inc & 1));
}

// ---------------------------------------------------------------------------
// Courant critical time step: per-element L/c then a block minimum (wave
// shuffles + LDS), one partial per block; k_cycle_end reduces the partials.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ double element_char_length(CJobDev *C, int64_t e) {
  const int64_t N = C->N;
  if constexpr (DIM == 2) {
    const DNLG_GLOBAL int32_t *cn = gbl(C->conn) + e * 4;
    const int n0 = cn[0], n1 = cn[1], n2 = cn[2], n3 = cn[3];
    // ElQua4NAx::getCharacteristicLength (ElQua4NAx.C:312-321)
    const double x31 = gbl(C->x)[n2] - gbl(C->x)[n0];
    const double x24 = gbl(C->x)[n1] - gbl(C->x)[n3];
    const double y42 = gbl(C->x)[N + n3] - gbl(C->x)[N + n1];
    const double y31 = gbl(C->x)[N + n2] - gbl(C->x)[N + n0];
    return ((x31 * y42 + x24 * y31) / sqrt(2.0 * (sqd(x24) + sqd(y42) + sqd(x31) + sqd(y31))));
  } else {
    // ElHex8N3D::getCharacteristicLength / getVolume (ElHex8N3D.C:614-684)
    const DNLG_GLOBAL int32_t *cn = gbl(C->conn) + e * 8;
    double X[8][3];
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int nd = cn[a];
#pragma unroll
      for (int c = 0; c < 3; c++) X[a][c] = gbl(C->x)[c * N + nd];
    }
    auto aert = [&](int p1, int p2, int p3, int p4) {
      const double x1 = X[p1][0], x2 = X[p2][0], x3 = X[p3][0], x4 = X[p4][0];
      const double y1 = X[p1][1], y2 = X[p2][1], y3 = X[p3][1], y4 = X[p4][1];
      const double z1 = X[p1][2], z2 = X[p2][2], z3 = X[p3][2], z4 = X[p4][2];
      return sqrt(4 * (sqd(x1 - x3) + sqd(y1 - y3) + sqd(z1 - z3)) * (sqd(x2 - x4) + sqd(y2 - y4) + sqd(z2 - z4)) -
                  4 * sqd(-(x2 * x3) + x1 * (x2 - x4) + x3 * x4 + y1 * y2 - y2 * y3 - y1 * y4 + y3 * y4 + z1 * z2 -
                          z2 * z3 - z1 * z4 + z3 * z4)) /
             4;
    };
    double A1 = aert(0, 1, 2, 3), A2 = aert(0, 1, 5, 4), A3 = aert(1, 2, 6, 5);
    double A4 = aert(3, 2, 6, 7), A5 = aert(0, 3, 7, 4), A6 = aert(4, 5, 6, 7);
    A1 = fmax(fmax(A1, A2), A3);
    A2 = fmax(fmax(A4, A5), A6);
    A1 = fmax(A1, A2);
    auto pyramid = [&](int o, int a, int b, int c, int apex) {
      double va[3], vb[3], vc[3], h[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        va[k] = X[a][k] - X[o][k];
        vb[k] = X[b][k] - X[o][k];
        vc[k] = X[c][k] - X[o][k];
        h[k] = X[apex][k] - X[o][k];
      }
      const double n1x = va[1] * vb[2] - va[2] * vb[1], n1y = va[2] * vb[0] - va[0] * vb[2],
                   n1z = va[0] * vb[1] - va[1] * vb[0];
      const double n2x = vc[1] * vb[2] - vc[2] * vb[1], n2y = vc[2] * vb[0] - vc[0] * vb[2],
                   n2z = vc[0] * vb[1] - vc[1] * vb[0];
      const double base = (sqrt(n1x * n1x + n1y * n1y + n1z * n1z) + sqrt(n2x * n2x + n2y * n2y + n2z * n2z)) / 2.0;
      double nx = va[1] * vc[2] - va[2] * vc[1], ny = va[2] * vc[0] - va[0] * vc[2], nz = va[0] * vc[1] - va[1] * vc[0];
      const double ln = sqrt(nx * nx + ny * ny + nz * nz);
      nx /= ln;
      ny /= ln;
      nz /= ln;
      const double ht = fabs(nx * h[0] + ny * h[1] + nz * h[2]);
      return base * ht / 3.0;
    };
    double vol = 0.0;
    vol += pyramid(0, 1, 2, 3, 7);
    vol += pyramid(0, 1, 5, 4, 7);
    vol += pyramid(1, 2, 6, 5, 7);
    return vol / A1;
  }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_courant(JobDev *const *jobs, const Task *tasks, int cs, int force) {
  __shared__ double s_min[4];
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  if (C->dt_method != DNLG_DT_COURANT) return;
  if (!force) {
    // Explicit.C:243-251: only while t_{n+1} < t_up_to, every dt_freq increments
    const Control c = load_control(C, cs);
    const int st0 = C->sync.st[0], st1 = C->sync.st[1];
    if (!c.running || ((c.inc & 1) ? st1 : st0) != 0 || !c.courant_due) return;
    if (!(c.t + c.dt < c.up_to)) return;
  }
  const int64_t e = tk.first + threadIdx.x;
  double v = 1.0e300;
  if (e < tk.end) {
    const int nip = C->nip;
    double rho = 0.0;
    for (int q = 0; q < nip; q++) rho += gbl(C->rho)[e * nip + q];
    rho /= nip;
    const double nu = C->nu;
    // Element::getElongationWaveSpeed (Element.C:343-359)
    const double c = sqrt((C->young * (1.0 - nu)) / (rho * (1.0 + nu) * (1.0 - 2.0 * nu)));
    v = element_char_length<DIM>(C, e) / c;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = s_min[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmin(m, s_min[w]);
    gbl(C->cmin)[tk.blk] = m;
  }
}

// ---------------------------------------------------------------------------
// Cycle step kernel.
// ---------------------------------------------------------------------------
// BC predictor on field0 of one node (the field0 half of
// Explicit::computePredictions, Explicit.C:468-480): constant BCs at the
// cycle's time and step, read by the element pass.
template <int DIM>
__device__ __forceinline__ void bc_field0(CJobDev *C, int64_t n, int f0, double t, double h) {
  const int64_t N = C->N;
  double ua[DIM], va[DIM], aa[DIM], ub[DIM], vb[DIM], ab[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    ua[d] = gbl(C->fu[f0])[d * N + n];
    va[d] = gbl(C->fv[f0])[d * N + n];
    aa[d] = gbl(C->fa[f0])[d * N + n];
    ub[d] = vb[d] = ab[d] = 0.0;
  }
  apply_bcs<DIM>(C, n, t, h, ua, va, aa, ub, vb, ab);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    gbl(C->fu[f0])[d * N + n] = ua[d];
    gbl(C->fv[f0])[d * N + n] = va[d];
    gbl(C->fa[f0])[d * N + n] = aa[d];
  }
}

// Minimum of the job's per-block Courant partials, over the whole block.
__device__ __forceinline__ double block_min_partials(CJobDev *C) {
  __shared__ double s_red[4];
  double v = 1.0e300;
  for (int k = threadIdx.x; k < C->n_cmin; k += blockDim.x) v = fmin(v, gbl(C->cmin)[k]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = s_red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmin(m, s_red[w]);
  return m;
}

// One launch between two cycles, per job: close the cycle that ran last, if
// any (with the Courant time step when it was due: the minimum over the
// per-block partials of k_courant), read from slot cs, and, with open != 0,
// open the next one and write the field0 BCs of the job's BC nodes.  The
// job's block 0 stores the new record in slot `out`.  Open: out = 1 - cs,
// max(1, ceil(n_bcn / 256)) blocks per job.  Close only: out = 0, one block
// per job (out may be cs: the block's waves have all read it first).
__global__ __launch_bounds__(256) void k_cycle_step(JobDev *const *jobs, const Task *tasks, int cs, int open,
                                                    int out) {
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  // through the global space: a close launch may store into slot cs
  Control c = load_control<const DNLG_GLOBAL Control>(C, cs);
  const int status = job_status(C);
  double freq = 0.0;
  bool stall = false;
  if (courant_ran(c, status)) {
    if (C->dt_method == DNLG_DT_POWER_ITERATION) {
      // the iteration slots of the launch sequence did not converge: close
      // the cycle and stop the job until the host-driven iteration finishes
      if (C->pi.done) freq = sqrt(C->pi.fmax);
      else stall = true;
    } else {
      freq = 2.0 / block_min_partials(C);
    }
  }
  const bool bc_work = close_open(C, c, status, freq, open != 0 && !stall);
  const int64_t k = tk.first + threadIdx.x;
  if (c.running && bc_work && k < tk.end) {
    const int64_t n = gbl(C->bcn)[k];
    if (C->dim == 3)
      bc_field0<3>(C, n, c.f0, c.t, c.dt);
    else
      bc_field0<2>(C, n, c.f0, c.t, c.dt);
  }
  __syncthreads();
  if (tk.blk == 0 && threadIdx.x == 0) {
    store_control(J, out, c);
    if (stall) gbl(&J->pi)->stall = 1;
  }
}

// ---------------------------------------------------------------------------
// Power-iteration time step (Model::computePowerIterationTimeStep,
// Model.C:942-1009).  One iteration = k_pi_elem (K_e v per element into its
// incidence slots), k_pi_node (v <- M^-1 (v + sum_e K_e v_e), block maxima of
// |v|), k_pi_step (fmax, convergence, buffer swap).  A job iterates while
// its pi.call is set and pi.done is not.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pi_active(CJobDev *C) {
  return C->dt_method == DNLG_DT_POWER_ITERATION && C->pi.call && !C->pi.done;
}

// Starts a time-step computation: force (the pre-loop's forced step,
// Explicit.C:152-154) or the cycle in slot cs ran and its step is due
// (courant_ran).  A job whose previous computation stalled keeps its state
// for the host loop.  One thread per job.
__global__ void k_pi_begin(JobDev *const *jobs, int n_jobs, int cs, int force) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  JobDev *J = jobs[j];
  if (J->dt_method != DNLG_DT_POWER_ITERATION || J->pi.stall) return;
  const Control &c = J->c[cs];
  const int status = J->sync.st[0] != 0 ? J->sync.st[0] : J->sync.st[1];
  const bool need = force ? status == 0 : courant_ran(c, status);
  J->pi.call = need ? 1 : 0;
  J->pi.done = 0;
  J->pi.iter = 0;
}

// K_e v matrix-free: per IP the strain B v, the stress C (B v) with the
// Hooke matrix of Material.C:98-152 (shear diagonal 1 - 2 nu, axisymmetric
// hoop row), and B^T C B v w detJ (x 2 pi r), on the Jacobian of the current
// coordinates (Explicit.C:245-249 computes it before the time step).  Equal
// to the reference's assembled K_e times v up to rounding.
template <int NN, int DIM, bool AXI>
__global__ __launch_bounds__(256) void k_pi_elem(JobDev *const *jobs, const Task *tasks) {
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  if (!pi_active(C)) return;
  const int64_t e = tk.first + threadIdx.x;
  if (e >= tk.end) return;
  const int64_t N = C->N;
  const DNLG_GLOBAL double *ev = gbl(C->pi_ev[C->pi.cur]);
  const double scale = C->pi.scale;
  double X[NN][DIM], v[NN][DIM], f[NN][DIM];
#pragma unroll
  for (int a = 0; a < NN; a++) {
    const int nd = gbl(C->conn)[e * NN + a];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      X[a][d] = gbl(C->x)[d * N + nd];
      v[a][d] = ev[d * N + nd] / scale;
      f[a][d] = 0.0;
    }
  }
  const double nu = C->nu;
  const double fct = C->young / ((1 + nu) * (1 - 2 * nu));
  const double c11 = (1 - nu) * fct, c12 = nu * fct, c44 = (1 - 2 * nu) * fct;
  for (int p = 0; p < NN; p++) {
    double Jm[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int j = 0; j < DIM; j++) Jm[i][j] = 0.0;
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) Jm[i][j] += tabD<NN>(p, a, j) * X[a][i];
    double det, inv[DIM][DIM];
    if constexpr (DIM == 3) {
      const double *dd = &Jm[0][0];
      det = dd[1] * dd[5] * dd[6] - dd[2] * dd[4] * dd[6] + dd[2] * dd[3] * dd[7] - dd[0] * dd[5] * dd[7] -
            dd[1] * dd[3] * dd[8] + dd[0] * dd[4] * dd[8];
      const double id = 1.0 / det;
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
      det = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
      const double id = 1.0 / det;
      inv[0][0] = Jm[1][1] * id;
      inv[0][1] = -Jm[0][1] * id;
      inv[1][0] = -Jm[1][0] * id;
      inv[1][1] = Jm[0][0] * id;
    }
    double dN[NN][DIM];
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int j = 0; j < DIM; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < DIM; k++) s += tabD<NN>(p, a, k) * inv[k][j];
        dN[a][j] = s;
      }
    if constexpr (DIM == 3) {
      double exx = 0, eyy = 0, ezz = 0, gxy = 0, gyz = 0, gxz = 0;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        exx += dN[a][0] * v[a][0];
        eyy += dN[a][1] * v[a][1];
        ezz += dN[a][2] * v[a][2];
        gxy += dN[a][1] * v[a][0] + dN[a][0] * v[a][1];
        gyz += dN[a][2] * v[a][1] + dN[a][1] * v[a][2];
        gxz += dN[a][2] * v[a][0] + dN[a][0] * v[a][2];
      }
      const double sxx = c11 * exx + c12 * (eyy + ezz), syy = c11 * eyy + c12 * (exx + ezz);
      const double szz = c11 * ezz + c12 * (exx + eyy);
      const double sxy = c44 * gxy, syz = c44 * gyz, sxz = c44 * gxz;
      const double w = det;  // Gauss weights are 1
#pragma unroll
      for (int a = 0; a < NN; a++) {
        f[a][0] += (dN[a][0] * sxx + dN[a][1] * sxy + dN[a][2] * sxz) * w;
        f[a][1] += (dN[a][1] * syy + dN[a][0] * sxy + dN[a][2] * syz) * w;
        f[a][2] += (dN[a][2] * szz + dN[a][1] * syz + dN[a][0] * sxz) * w;
      }
    } else if constexpr (!AXI) {
      // plane strain (Material.C:104-111): the leading 3 x 3 block, unit thickness
      double exx = 0, eyy = 0, gxy = 0;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        exx += dN[a][0] * v[a][0];
        eyy += dN[a][1] * v[a][1];
        gxy += dN[a][1] * v[a][0] + dN[a][0] * v[a][1];
      }
      const double sxx = c11 * exx + c12 * eyy, syy = c11 * eyy + c12 * exx, sxy = c44 * gxy;
      const double w = det;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        f[a][0] += (dN[a][0] * sxx + dN[a][1] * sxy) * w;
        f[a][1] += (dN[a][1] * syy + dN[a][0] * sxy) * w;
      }
    } else {
      double r = 0.0;
#pragma unroll
      for (int a = 0; a < NN; a++) r += tabN<NN>(p, a) * X[a][0];
      double err = 0, ezz = 0, grz = 0, ett = 0;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        err += dN[a][0] * v[a][0];
        ezz += dN[a][1] * v[a][1];
        grz += dN[a][1] * v[a][0] + dN[a][0] * v[a][1];
        ett += tabN<NN>(p, a) / r * v[a][0];
      }
      const double srr = c11 * err + c12 * (ezz + ett), szz = c11 * ezz + c12 * (err + ett);
      const double stt = c11 * ett + c12 * (err + ezz), srz = c44 * grz;
      const double w = det * k2PI * r;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        f[a][0] += (dN[a][0] * srr + dN[a][1] * srz + tabN<NN>(p, a) / r * stt) * w;
        f[a][1] += (dN[a][1] * szz + dN[a][0] * srz) * w;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < NN; a++) {
    DNLG_GLOBAL double *out = gbl(C->elemF) + (int64_t)gbl(C->e2slot)[e * NN + a] * DIM;
#pragma unroll
    for (int d = 0; d < DIM; d++) out[d] = f[a][d];
  }
}

// v <- M^-1 (v + sum_e K_e v_e): the element products summed per node in
// element order onto the normalised iterate (Vector::gatherFrom,
// Vector.C:893-945), divided by the lumped mass (MatrixDiag::divideBy,
// MatrixDiag.C:655-670); the block's maximum |v| to pi_bmax.
template <int DIM>
__global__ __launch_bounds__(kNodeBlock) void k_pi_node(JobDev *const *jobs, const Task *tasks) {
  __shared__ double s_max[kNodeBlock / 64];
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  if (!pi_active(C)) return;
  const int64_t n = tk.first + threadIdx.x;
  double vmax = 0.0;
  if (n < tk.end) {
    const int64_t N = C->N;
    const int cur = C->pi.cur;
    const DNLG_GLOBAL double *src = gbl(C->pi_ev[cur]);
    DNLG_GLOBAL double *dst = gbl(C->pi_ev[1 - cur]);
    const double scale = C->pi.scale;
    double acc[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) acc[d] = src[d * N + n] / scale;
    const DNLG_GLOBAL double *ef = gbl(C->elemF);
    const int64_t k1 = gbl(C->n2e_ptr)[n + 1];
    int64_t k = gbl(C->n2e_ptr)[n];
    for (; k + 4 <= k1; k += 4) add_slots<4, DIM>(ef, k, acc);
    if (k + 2 <= k1) {
      add_slots<2, DIM>(ef, k, acc);
      k += 2;
    }
    if (k < k1) add_slots<1, DIM>(ef, k, acc);
    const double m = gbl(C->mass)[n];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      acc[d] /= m;
      dst[d * N + n] = acc[d];
      vmax = fmax(vmax, fabs(acc[d]));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, off));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = vmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = s_max[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, s_max[w]);
    gbl(C->pi_bmax)[tk.blk] = m;
  }
}

// fmax = max |v|, convergence |fmax - fmax_prev| / fmax < 1e-4, at most
// 1000 iterations, checked in the reference's order (Model.C:990-1006);
// the new iterate is normalised by fmax where it is read.  One block per
// job; cs: the control slot whose increment parity takes a failure.
__global__ __launch_bounds__(256) void k_pi_step(JobDev *const *jobs, const Task *tasks, int cs) {
  __shared__ double s_max[4];
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  if (!pi_active(C)) return;
  double v = 0.0;
  for (int k = threadIdx.x; k < C->n_pibmax; k += blockDim.x) v = fmax(v, gbl(C->pi_bmax)[k]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double fmax_ = s_max[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) fmax_ = fmax(fmax_, s_max[w]);
    DNLG_GLOBAL PiState &P = *gbl(&J->pi);
    const double convergence = fabs(fmax_ - P.fmax) / fmax_;
    P.fmax = fmax_;
    P.scale = fmax_;
    P.cur = 1 - P.cur;
    P.iter = P.iter + 1;
    if (P.iter > 1000) {
      set_status(J, DNLG_E_PI_DIVERGED, (int)(C->c[cs].inc & 1));
      P.done = 1;
    } else if (convergence < 1e-4) {
      P.done = 1;
    }
  }
}

// After a host-driven iteration: the jobs it ran for (preloop != 0: every
// job that iterated in the pre-loop; otherwise the stalled jobs) take their
// step, safety Omega_s / sqrt(fmax), in slot 0 when the iteration
// converged; the stall and call flags are cleared.  One thread per job.
__global__ void k_pi_finish(JobDev *const *jobs, int n_jobs, int preloop) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  JobDev *J = jobs[j];
  if (J->dt_method != DNLG_DT_POWER_ITERATION || !J->pi.call) return;
  if (!preloop && !J->pi.stall) return;
  if (J->pi.done && J->sync.st[0] == 0 && J->sync.st[1] == 0) {
    const double freq = sqrt(J->pi.fmax);
    J->c[0].dt = J->safety * J->omegaS / freq;
  }
  J->pi.call = 0;
  J->pi.stall = 0;
}

// Jobs of the batch still iterating: out[j] = 1 (host-driven loops).
__global__ void k_pi_pending(JobDev *const *jobs, int n_jobs, int *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  const JobDev *J = jobs[j];
  out[j] = (J->dt_method == DNLG_DT_POWER_ITERATION && J->pi.call && !J->pi.done) ? 1 : 0;
}

// Host-loop snapshot: t, dt, up_to, status, increment, power-iteration stall
// of every job.
__global__ void k_snapshot(JobDev *const *jobs, int n_jobs, double *out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  const JobDev *J = jobs[j];
  const Control &c = J->c[0];
  double *o = out + kSnapDoubles * j;
  o[0] = c.t;
  o[1] = c.dt;
  o[2] = c.up_to;
  o[3] = (double)(J->sync.st[0] != 0 ? J->sync.st[0] : J->sync.st[1]);
  o[4] = (double)c.inc;
  o[5] = (double)J->pi.stall;
  o[6] = (double)J->res.hold;
}

// Target time of the current advance, per job.
__global__ void k_set_up_to(JobDev *const *jobs, int n_jobs, const double *up_to) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  jobs[j]->c[0].up_to = up_to[j];
}

// Forced time step of the pre-loop (Explicit.C:152-154).
__global__ void k_force_dt(JobDev *const *jobs, int n_jobs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  JobDev *J = jobs[j];
  if (J->dt_method != DNLG_DT_COURANT) return;  // k_pi_finish sets it
  double m = J->cmin[0];
  for (int k = 1; k < J->n_cmin; k++) m = fmin(m, J->cmin[k]);
  const double fmax = 2.0 / m;
  J->c[0].dt = J->safety * J->omegaS / fmax;
}

// ---------------------------------------------------------------------------
// Pre-loop: reference Jacobian -> detJ0 (Element3D.C:112-113,
// ElementPlane.C:92-100) and the element lumped masses
// (Element::computeMassMatrix, Element.C:304-339) into elemF's first slot.
// ---------------------------------------------------------------------------
template <int NN, int DIM, bool AXI>
__global__ __launch_bounds__(256) void k_prep(JobDev *const *jobs, const Task *tasks) {
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  const int64_t e = tk.first + threadIdx.x;
  if (e >= tk.end) return;
  const int64_t N = J->N;
  double X[NN][DIM];
#pragma unroll
  for (int a = 0; a < NN; a++) {
    const int nd = J->conn[e * NN + a];
#pragma unroll
    for (int d = 0; d < DIM; d++) X[a][d] = J->x[d * N + nd];
  }
  double me[NN];
#pragma unroll
  for (int a = 0; a < NN; a++) me[a] = 0.0;
  bool neg = false;
  for (int p = 0; p < NN; p++) {
    double Jm[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int jj = 0; jj < DIM; jj++) Jm[i][jj] = 0.0;
#pragma unroll
    for (int a = 0; a < NN; a++)
#pragma unroll
      for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int jj = 0; jj < DIM; jj++) Jm[i][jj] += tabD<NN>(p, a, jj) * X[a][i];
    double det;
    if constexpr (DIM == 3) {
      const double *dd = &Jm[0][0];
      det = dd[1] * dd[5] * dd[6] - dd[2] * dd[4] * dd[6] + dd[2] * dd[3] * dd[7] - dd[0] * dd[5] * dd[7] -
            dd[1] * dd[3] * dd[8] + dd[0] * dd[4] * dd[8];
    } else {
      det = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
    }
    neg = neg || (det < 0.0);
    double wdj = det;
    double det0 = det;
    if constexpr (AXI) {
      double r = 0.0;
#pragma unroll
      for (int a = 0; a < NN; a++) r += tabN<NN>(p, a) * X[a][0];
      det0 = det * r;
      wdj *= k2PI * r;
    }
    J->detJ0[e * NN + p] = det0;
    if (J->detJ0f) J->detJ0f[e * NN + p] = (float)det0;
#pragma unroll
    for (int a = 0; a < NN; a++) me[a] += J->rho0 * tabN<NN>(p, a) * wdj;
  }
  if (neg) set_status(J, DNLG_E_NEG_JACOBIAN, 0);
#pragma unroll
  for (int a = 0; a < NN; a++) J->elemF[(int64_t)J->e2slot[e * NN + a] * DIM] = me[a];
}

// Model::computeMassMatrix assembly (Model.C:537-590), gathered per node in
// element order.
template <int DIM>
__global__ __launch_bounds__(kNodeBlock) void k_mass_gather(JobDev *const *jobs, const Task *tasks) {
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  const int64_t n = tk.first + threadIdx.x;
  if (n >= tk.end) return;
  double m = 0.0;
  for (int64_t k = J->n2e_ptr[n]; k < J->n2e_ptr[n + 1]; k++) m += J->elemF[k * DIM];
  J->mass[n] = m;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
void upload_tables(const double *hexN, const double *hexD, const double *quaN, const double *quaD) {
  hipMemcpyToSymbol(HIP_SYMBOL(c_hexN), hexN, sizeof(double) * 64);
  hipMemcpyToSymbol(HIP_SYMBOL(c_hexD), hexD, sizeof(double) * 192);
  hipMemcpyToSymbol(HIP_SYMBOL(c_quaN), quaN, sizeof(double) * 16);
  hipMemcpyToSymbol(HIP_SYMBOL(c_quaD), quaD, sizeof(double) * 32);
  float f[192];
  for (int k = 0; k < 64; k++) f[k] = (float)hexN[k];
  hipMemcpyToSymbol(HIP_SYMBOL(c_hexNf), f, sizeof(float) * 64);
  for (int k = 0; k < 192; k++) f[k] = (float)hexD[k];
  hipMemcpyToSymbol(HIP_SYMBOL(c_hexDf), f, sizeof(float) * 192);
  for (int k = 0; k < 16; k++) f[k] = (float)quaN[k];
  hipMemcpyToSymbol(HIP_SYMBOL(c_quaNf), f, sizeof(float) * 16);
  for (int k = 0; k < 32; k++) f[k] = (float)quaD[k];
  hipMemcpyToSymbol(HIP_SYMBOL(c_quaDf), f, sizeof(float) * 32);
}

int elem_block_threads(int type) { return type == DNLG_EL_HEX8N_3D ? 128 : 256; }
int elem_block_threads_mixed(int type) { return type == DNLG_EL_HEX8N_3D ? 128 : 256; }

// One launch of an element kernel template: with events (hipExtLaunchKernelGGL
// records start / stop on the stream) or without.
template <typename K>
void launch_elem(K kern, int block, int nblk, hipStream_t s, const hipEvent_t *ev, JobDev *const *jobs,
                 const Task *tasks, int cs, int derive) {
  if (ev)
    hipExtLaunchKernelGGL(kern, dim3(nblk), dim3(block), 0, s, ev[0], ev[1], 0, jobs, tasks, cs, derive);
  else
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(block), 0, s, jobs, tasks, cs, derive);
}

template <bool RR64>
void launch_elem_mixed_rr(int type, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                          hipStream_t s, const hipEvent_t *ev) {
  if (type == DNLG_EL_HEX8N_3D)
    launch_elem(k_elem_mixed<8, 3, 128, false, RR64>, 128, nblk, s, ev, jobs, tasks, cs, derive);
  else if (type == DNLG_EL_QUA4N_AX)
    launch_elem(k_elem_mixed<4, 2, 256, true, RR64>, 256, nblk, s, ev, jobs, tasks, cs, derive);
  else
    launch_elem(k_elem_mixed<4, 2, 256, false, RR64>, 256, nblk, s, ev, jobs, tasks, cs, derive);
}

void launch_elem_mixed(int type, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                       hipStream_t s, const hipEvent_t *ev, bool rr64) {
  if (nblk <= 0) return;
  if (rr64)
    launch_elem_mixed_rr<true>(type, jobs, tasks, nblk, cs, derive, s, ev);
  else
    launch_elem_mixed_rr<false>(type, jobs, tasks, nblk, cs, derive, s, ev);
}

template <bool ANN>
void launch_elem_fp64_law(int type, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                          hipStream_t s, const hipEvent_t *ev) {
  if (type == DNLG_EL_HEX8N_3D)
    launch_elem(k_elem_fp64<8, 3, 128, false, ANN>, 128, nblk, s, ev, jobs, tasks, cs, derive);
  else if (type == DNLG_EL_QUA4N_AX)
    launch_elem(k_elem_fp64<4, 2, 256, true, ANN>, 256, nblk, s, ev, jobs, tasks, cs, derive);
  else
    launch_elem(k_elem_fp64<4, 2, 256, false, ANN>, 256, nblk, s, ev, jobs, tasks, cs, derive);
}

void launch_elem_fp64(int type, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                      hipStream_t s, const hipEvent_t *ev, bool ann) {
  if (nblk <= 0) return;
  if (ann)
    launch_elem_fp64_law<true>(type, jobs, tasks, nblk, cs, derive, s, ev);
  else
    launch_elem_fp64_law<false>(type, jobs, tasks, nblk, cs, derive, s, ev);
}

void launch_correct(int dim, bool f32, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int fold,
                    hipStream_t s) {
  if (nblk <= 0) return;
  if (dim == 3) {
    if (f32)
      hipLaunchKernelGGL((k_correct<3, float>), dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks, cs, fold);
    else
      hipLaunchKernelGGL((k_correct<3, double>), dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks, cs, fold);
  } else {
    if (f32)
      hipLaunchKernelGGL((k_correct<2, float>), dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks, cs, fold);
    else
      hipLaunchKernelGGL((k_correct<2, double>), dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks, cs, fold);
  }
}

void launch_courant(int dim, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int force, hipStream_t s) {
  if (nblk <= 0) return;
  if (dim == 3)
    hipLaunchKernelGGL(k_courant<3>, dim3(nblk), dim3(256), 0, s, jobs, tasks, cs, force);
  else
    hipLaunchKernelGGL(k_courant<2>, dim3(nblk), dim3(256), 0, s, jobs, tasks, cs, force);
}

void launch_cycle_step(JobDev *const *jobs, const Task *tasks, int nblk, int cs, int open, int out, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_cycle_step, dim3(nblk), dim3(256), 0, s, jobs, tasks, cs, open, out);
}
void launch_pi_begin(JobDev *const *jobs, int n, int cs, int force, hipStream_t s) {
  hipLaunchKernelGGL(k_pi_begin, dim3((n + 63) / 64), dim3(64), 0, s, jobs, n, cs, force);
}
void launch_pi_elem(int type, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s) {
  if (nblk <= 0) return;
  if (type == DNLG_EL_HEX8N_3D)
    hipLaunchKernelGGL((k_pi_elem<8, 3, false>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
  else if (type == DNLG_EL_QUA4N_AX)
    hipLaunchKernelGGL((k_pi_elem<4, 2, true>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
  else
    hipLaunchKernelGGL((k_pi_elem<4, 2, false>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
}
void launch_pi_node(int dim, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s) {
  if (nblk <= 0) return;
  if (dim == 3)
    hipLaunchKernelGGL(k_pi_node<3>, dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks);
  else
    hipLaunchKernelGGL(k_pi_node<2>, dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks);
}
void launch_pi_step(JobDev *const *jobs, const Task *tasks, int nblk, int cs, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_pi_step, dim3(nblk), dim3(256), 0, s, jobs, tasks, cs);
}
void launch_pi_finish(JobDev *const *jobs, int n, int preloop, hipStream_t s) {
  hipLaunchKernelGGL(k_pi_finish, dim3((n + 63) / 64), dim3(64), 0, s, jobs, n, preloop);
}
void launch_pi_pending(JobDev *const *jobs, int n, int *out, hipStream_t s) {
  hipLaunchKernelGGL(k_pi_pending, dim3((n + 63) / 64), dim3(64), 0, s, jobs, n, out);
}
void launch_snapshot(JobDev *const *jobs, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_snapshot, dim3((n + 63) / 64), dim3(64), 0, s, jobs, n, out);
}
void launch_set_up_to(JobDev *const *jobs, int n, const double *up_to, hipStream_t s) {
  hipLaunchKernelGGL(k_set_up_to, dim3((n + 63) / 64), dim3(64), 0, s, jobs, n, up_to);
}
void launch_force_dt(JobDev *const *jobs, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_force_dt, dim3((n + 63) / 64), dim3(64), 0, s, jobs, n);
}
void launch_prep(int type, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s) {
  if (nblk <= 0) return;
  if (type == DNLG_EL_HEX8N_3D)
    hipLaunchKernelGGL((k_prep<8, 3, false>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
  else if (type == DNLG_EL_QUA4N_AX)
    hipLaunchKernelGGL((k_prep<4, 2, true>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
  else
    hipLaunchKernelGGL((k_prep<4, 2, false>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
}
void launch_mass_gather(int dim, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s) {
  if (nblk <= 0) return;
  if (dim == 3)
    hipLaunchKernelGGL(k_mass_gather<3>, dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks);
  else
    hipLaunchKernelGGL(k_mass_gather<2>, dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks);
}

}  // namespace dnlg
