// elem_fp64.hip -- the FP64 element kernel (strains, pressure, radial
// return, rotation, internal forces, densities of one cycle) and its
// launchers, one instantiation per schedule (sched.h).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "control.h"
#include "kernels.h"
#include "material.h"
#include "polar_fp64.h"
#include "sched.h"
#include "shape.h"

namespace dnlg {

// ---------------------------------------------------------------------------
// Element kernel, FP64.  One lane per integration point (NIP lanes per
// element, NIP == NN for the hex and quads); the element's nodes are staged
// in LDS by its own lanes, per-IP inverse Jacobians, stresses and weights go back to
// LDS so that lane a can sum node a's force over the IPs in the reference's
// order (IP, then direction j).
// ---------------------------------------------------------------------------
// Three waves per SIMD (168 VGPRs, a short spill); the LDS of a 16-hex block
// (24 KB) allows three.  Measured 11 % faster than two waves without spill.
// ANN: the variant for jobs whose law is DNLG_LAW_ANN2SIG (its network
// evaluation keeps per-lane arrays that the other laws' variants do without).
// TET (NN = 4, DIM = 3): the 4-node tetrahedron of ElTet4N3D.C, one
// integration point (ElTet4N3D.C:62-71: the centroid, weight 1/6, linear
// shape functions): four lanes per element stage its nodes and sum its
// nodal forces; lane 0 computes the integration point.
template <int NN, int DIM, int BLOCK, bool AXI, bool ANN, int WAVES, bool TET = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_elem_fp64(
    JobDev *const *jobs, const Task *tasks, int cs, int derive) {
  constexpr int NIP = TET ? 1 : NN;  // integration points per element
  constexpr int EPB = BLOCK / NN;    // elements per block (NN lanes each)
  __shared__ double s_x[EPB][NN][DIM];
  __shared__ double s_u[EPB][NN][DIM];
  __shared__ double s_inv[EPB][NIP][DIM][DIM];  // J^-1 per IP, for the force phase
  __shared__ double s_sw[EPB][NIP][8];
  __shared__ double s_tr[EPB][NIP];

  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  // The job-record fields of the control derivation and of the node loads
  // in one batch of scalar loads ahead of the first branch (the empty asm
  // takes them as operands, so none is sunk behind a branch to cost a round
  // trip of its own; cycle_control reads the same constant-space words).
  const int64_t N = C->N;
  const DNLG_GLOBAL int32_t *conn = gbl(C->conn);
  const DNLG_GLOBAL double *xg = gbl(C->x), *fu0 = gbl(C->fu[0]), *fu1 = gbl(C->fu[1]);
  {
    const Control c0 = load_control(C, cs);
    asm volatile("" ::"s"(tk.first), "s"(tk.end), "s"(tk.blk), "s"(c0.t), "s"(c0.dt), "s"(c0.up_to), "s"(c0.inc),
                 "s"(c0.f0), "s"(c0.running), "s"(c0.pending), "s"(C->sync.st[0]), "s"(C->sync.st[1]),
                 "s"(C->pi.stall), "s"(C->res.hold), "s"(C->dt_freq), "s"(C->density_always), "s"(N), "s"(conn),
                 "s"(xg), "s"(fu0), "s"(fu1));
  }
  const Control cc = cycle_control(J, C, cs, derive, tk.blk);
  if (!cc.running) return;
  const int par = (int)(cc.inc & 1);
  const int le = threadIdx.x / NN;
  const int p = threadIdx.x % NN;
  const int64_t e = tk.first + le;
  const bool valid = e < tk.end;
  const bool ipl = valid && p < NIP;  // this lane computes integration point p
  const int f0 = cc.f0;

  // stage the element's nodes: lane p brings node p
  if (valid) {
    const int nd = conn[e * NN + p];
    const DNLG_GLOBAL double *u0 = f0 ? fu1 : fu0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      s_x[le][p][d] = xg[d * N + nd];
      s_u[le][p][d] = u0[d * N + nd];
    }
  }
  __syncthreads();

  const int64_t G = C->G;
  const int64_t g = e * NIP + p;
  double dN[NN][DIM];
  double detJ = 1.0, radius = 1.0;
  double dE[6] = {0, 0, 0, 0, 0, 0}, R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (ipl) {
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
        for (int j = 0; j < DIM; j++) Jm[i][j] += shapeD<NN, TET>(p, a, j) * s_x[le][a][i];
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
        for (int k = 0; k < DIM; k++) s += shapeD<NN, TET>(p, a, k) * inv[k][j];
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
  if (ipl) {
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
    // Gauss weights: 1 (ElQua4NAx.C:63-110, ElHex8N3D.C:93-231), 1/6 for the
    // tetrahedron's point
    wdj = TET ? ip_weight<TET>() * detJ : detJ;
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
        for (int k = 0; k < DIM; k++) d += shapeD<NN, TET>(q, a, k) * s_inv[le][q][k][j];
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

// One schedule of the FP64 element kernel of one element type and law
// variant.
template <int NN, int DIM, bool AXI, bool ANN, int BLOCK>
void launch_fp64_block(int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                       hipStream_t s, const hipEvent_t *ev) {
  if (waves == kFP64WavesAlt)
    launch_elem(k_elem_fp64<NN, DIM, BLOCK, AXI, ANN, kFP64WavesAlt>, BLOCK, nblk, s, ev, jobs, tasks, cs, derive);
  else
    launch_elem(k_elem_fp64<NN, DIM, BLOCK, AXI, ANN, kFP64Waves>, BLOCK, nblk, s, ev, jobs, tasks, cs, derive);
}

template <int NN, int DIM, bool AXI, bool ANN>
void launch_fp64_type(int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                      hipStream_t s, const hipEvent_t *ev) {
  constexpr int B0 = NN == 8 ? 64 : 128;
  if (block == B0)
    launch_fp64_block<NN, DIM, AXI, ANN, B0>(waves, jobs, tasks, nblk, cs, derive, s, ev);
  else if (block == 4 * B0)
    launch_fp64_block<NN, DIM, AXI, ANN, 4 * B0>(waves, jobs, tasks, nblk, cs, derive, s, ev);
  else
    launch_fp64_block<NN, DIM, AXI, ANN, 2 * B0>(waves, jobs, tasks, nblk, cs, derive, s, ev);
}

void launch_elem_fp64(int type, int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs,
                      int derive, hipStream_t s, const hipEvent_t *ev, bool ann) {
  if (nblk <= 0) return;
  if (type == DNLG_EL_TET4N_3D) {  // default schedule only: 64 tetrahedra per block
    if (ann)
      launch_elem(k_elem_fp64<4, 3, 256, false, true, kFP64Waves, true>, 256, nblk, s, ev, jobs, tasks, cs, derive);
    else
      launch_elem(k_elem_fp64<4, 3, 256, false, false, kFP64Waves, true>, 256, nblk, s, ev, jobs, tasks, cs, derive);
    return;
  }
  if (ann) {  // the network-law variant: default schedule only
    if (type == DNLG_EL_HEX8N_3D)
      launch_elem(k_elem_fp64<8, 3, 128, false, true, kFP64Waves>, 128, nblk, s, ev, jobs, tasks, cs, derive);
    else if (type == DNLG_EL_QUA4N_AX)
      launch_elem(k_elem_fp64<4, 2, 256, true, true, kFP64Waves>, 256, nblk, s, ev, jobs, tasks, cs, derive);
    else
      launch_elem(k_elem_fp64<4, 2, 256, false, true, kFP64Waves>, 256, nblk, s, ev, jobs, tasks, cs, derive);
    return;
  }
  if (type == DNLG_EL_HEX8N_3D)
    launch_fp64_type<8, 3, false, false>(block, waves, jobs, tasks, nblk, cs, derive, s, ev);
  else if (type == DNLG_EL_QUA4N_AX)
    launch_fp64_type<4, 2, true, false>(block, waves, jobs, tasks, nblk, cs, derive, s, ev);
  else  // plane strain: default schedule only
    launch_elem(k_elem_fp64<4, 2, 256, false, false, kFP64Waves>, 256, nblk, s, ev, jobs, tasks, cs, derive);
}

void upload_tables_fp64(const double *hexN, const double *hexD, const double *quaN, const double *quaD) {
  upload_shape_tables(hexN, hexD, quaN, quaD);
}

}  // namespace dnlg
