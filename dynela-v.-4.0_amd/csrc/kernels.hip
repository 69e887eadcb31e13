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

#include "control.h"
#include "kernels.h"
#include "material.h"
#include "polar_fp64.h"
#include "shape.h"

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
// are read from its incidence slots in element order (same
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

// Adds node n's incident forces to f in element order: its slot_w fixed
// slots (those past its incidence count hold zeros, which change no sum),
// then its overflow slots, if any.  The fixed slots' addresses need no
// index load, so their loads issue with the node's own values.  MAXB: the
// slots loaded together.
template <int MAXB, int DIM, typename FT>
__device__ __forceinline__ void add_node_slots(const DNLG_GLOBAL FT *ef, int W, const DNLG_GLOBAL int32_t *ovf,
                                               int64_t N, int64_t n, double f[DIM]) {
  const int64_t k0 = n * W;
  int k = 0;
  for (; k + MAXB <= W; k += MAXB) add_slots<MAXB, DIM>(ef, k0 + k, f);
  if constexpr (MAXB > 4) {
    if (k + 4 <= W) {
      add_slots<4, DIM>(ef, k0 + k, f);
      k += 4;
    }
  }
  if constexpr (MAXB > 2) {
    if (k + 2 <= W) {
      add_slots<2, DIM>(ef, k0 + k, f);
      k += 2;
    }
  }
  if (k < W) add_slots<1, DIM>(ef, k0 + k, f);
  if (ovf) {
    const int64_t o1 = N * W + ovf[n + 1];
    for (int64_t o = N * W + ovf[n]; o < o1; o++) add_slots<1, DIM>(ef, o, f);
  }
}

// MAXB: the incidence slots loaded together (a schedule parameter; the
// sums run in slot order whatever the batching)
// Latency: with one node per lane the launch has about two waves per SIMD,
// so a wave's chain of dependent loads is the kernel's duration.  The task,
// the control record and every job-record field the node update needs are
// read as one batch of scalar loads ahead of the first branch (the empty asm
// takes them all as operands, so none of them is sunk behind a branch, where
// each would cost a round trip of its own), and the node's own values --
// coordinates included -- are loaded before the incidence slots.
template <int DIM, typename FT, int MAXB>
__global__ __launch_bounds__(kNodeBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_correct(
    JobDev *const *jobs, const Task *tasks, int cs, int fold) {
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  const Control c = load_control(C, cs);
  const int st0 = C->sync.st[0], st1 = C->sync.st[1];
  const int64_t N = C->N;
  const int W = C->slot_w;
  const DNLG_GLOBAL int32_t *ovf = gbl(C->ovf_ptr);
  const DNLG_GLOBAL int32_t *bc_ptr = gbl(C->bc_ptr);
  const DNLG_GLOBAL double *mass = gbl(C->mass);
  DNLG_GLOBAL double *x = gbl(C->x);
  DNLG_GLOBAL double *fv0 = gbl(C->fv[0]), *fv1 = gbl(C->fv[1]);
  DNLG_GLOBAL double *fa0 = gbl(C->fa[0]), *fa1 = gbl(C->fa[1]);
  const DNLG_GLOBAL FT *ef;
  if constexpr (sizeof(FT) == 8) ef = gbl(C->elemF); else ef = gbl(C->elemF32);
  const double alphaM = C->alphaM, gam = C->gammaCH, beta = C->beta;
  asm volatile("" ::"s"(tk.first), "s"(tk.end), "s"(c.dt), "s"(c.inc), "s"(c.f0), "s"(c.running), "s"(st0),
               "s"(st1), "s"(N), "s"(W), "s"(ovf), "s"(bc_ptr), "s"(mass), "s"(x), "s"(fv0), "s"(fv1), "s"(fa0),
               "s"(fa1), "s"(ef), "s"(alphaM), "s"(gam), "s"(beta));
  if (!c.running || ((c.inc & 1) ? st1 : st0) != 0) return;
  const int64_t n = tk.first + threadIdx.x;
  if (n >= tk.end) return;
  const int a = c.f0, b = 1 - a;
  const double h = c.dt;
  DNLG_GLOBAL double *fva = a ? fv1 : fv0, *faa = a ? fa1 : fa0;
  DNLG_GLOBAL double *fvb = a ? fv0 : fv1, *fab = a ? fa0 : fa1;
  // the node's own values, then the incidence slots in batches (the fixed
  // slots' loads issue with these)
  const bool has_bc = bc_ptr[n + 1] > bc_ptr[n];
  const double m = mass[n];
  double va[DIM], aa[DIM], xo[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    va[d] = fva[d * N + n];
    aa[d] = faa[d * N + n];
    xo[d] = x[d * N + n];
  }
  double f[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) f[d] = 0.0;
  add_node_slots<MAXB, DIM>(ef, W, ovf, N, n, f);
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
  if (has_bc) {
    DNLG_GLOBAL double *fua = gbl(C->fu[a]);
#pragma unroll
    for (int d = 0; d < DIM; d++) ua[d] = fua[d * N + n];
    apply_bcs<DIM>(C, n, c.t, h, ua, va, aa, ub, vb, ab);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      fua[d * N + n] = ua[d];
      fva[d * N + n] = va[d];
      faa[d * N + n] = aa[d];
    }
  }
  bool finite = true;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const double xn = xo[d] + ub[d];
    x[d * N + n] = xn;
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
  DNLG_GLOBAL double *fub = gbl(C->fu[b]);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    fub[d * N + n] = ub[d];
    fvb[d * N + n] = vb[d];
    fab[d * N + n] = ab[d];
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
    if (is_pi(C->dt_method)) {
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
  return is_pi(C->dt_method) && C->pi.call && !C->pi.done;
}

// Starts a time-step computation: force (the pre-loop's forced step,
// Explicit.C:152-154) or the cycle in slot cs ran and its step is due
// (courant_ran).  A job whose previous computation stalled keeps its state
// for the host loop.  One thread per job.
__global__ void k_pi_begin(JobDev *const *jobs, int n_jobs, int cs, int force) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  JobDev *J = jobs[j];
  if (!is_pi(J->dt_method) || J->pi.stall) return;
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
template <int NN, int DIM, bool AXI, bool TET = false>
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
  // PowerIterationUnder (computeElasticStiffnessMatrix(true)): one point at
  // the element centre, weight NN (8 / 4), shape functions 1 / NN and
  // derivatives +-1 / NN by the nodes' local signs (ElHex8N3D.C:233-252,
  // ElQua4NAx.C:112-124, ElQua4N2D.C:115-127), on the current coordinates
  // (Element3D::computeUnderJacobian, ElementPlane::computeUnderJacobian)
  // The tetrahedron's one point is also its under-integration point
  // (ElTet4N3D.C:73-82), with weight 1/6.
  const bool under = !TET && C->dt_method == DNLG_DT_POWER_ITERATION_UNDER;
  const int npts = (TET || under) ? 1 : NN;
  auto dshape = [&](int p, int a, int k) { return under ? node_sign<NN>(a, k) / NN : shapeD<NN, TET>(p, a, k); };
  auto shape = [&](int p, int a) { return under ? 1.0 / NN : shapeN<NN, TET>(p, a); };
  const double wgt = TET ? ip_weight<TET>() : (under ? (double)NN : 1.0);  // Gauss weights 1
  for (int p = 0; p < npts; p++) {
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
        for (int j = 0; j < DIM; j++) Jm[i][j] += dshape(p, a, j) * X[a][i];
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
        for (int k = 0; k < DIM; k++) s += dshape(p, a, k) * inv[k][j];
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
      const double w = wgt * det;
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
      const double w = wgt * det;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        f[a][0] += (dN[a][0] * sxx + dN[a][1] * sxy) * w;
        f[a][1] += (dN[a][1] * syy + dN[a][0] * sxy) * w;
      }
    } else {
      double r = 0.0;
#pragma unroll
      for (int a = 0; a < NN; a++) r += shape(p, a) * X[a][0];
      double err = 0, ezz = 0, grz = 0, ett = 0;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        err += dN[a][0] * v[a][0];
        ezz += dN[a][1] * v[a][1];
        grz += dN[a][1] * v[a][0] + dN[a][0] * v[a][1];
        ett += shape(p, a) / r * v[a][0];
      }
      const double srr = c11 * err + c12 * (ezz + ett), szz = c11 * ezz + c12 * (err + ett);
      const double stt = c11 * ett + c12 * (err + ezz), srz = c44 * grz;
      const double w = wgt * det * k2PI * r;
#pragma unroll
      for (int a = 0; a < NN; a++) {
        f[a][0] += (dN[a][0] * srr + dN[a][1] * srz + shape(p, a) / r * stt) * w;
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
    add_node_slots<4, DIM>(gbl(C->elemF), C->slot_w, gbl(C->ovf_ptr), N, n, acc);
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
  if (!is_pi(J->dt_method) || !J->pi.call) return;
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
  out[j] = (is_pi(J->dt_method) && J->pi.call && !J->pi.done) ? 1 : 0;
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
template <int NN, int DIM, bool AXI, bool TET = false>
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
  constexpr int NIP = TET ? 1 : NN;
  for (int p = 0; p < NIP; p++) {
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
        for (int jj = 0; jj < DIM; jj++) Jm[i][jj] += shapeD<NN, TET>(p, a, jj) * X[a][i];
    double det;
    if constexpr (DIM == 3) {
      const double *dd = &Jm[0][0];
      det = dd[1] * dd[5] * dd[6] - dd[2] * dd[4] * dd[6] + dd[2] * dd[3] * dd[7] - dd[0] * dd[5] * dd[7] -
            dd[1] * dd[3] * dd[8] + dd[0] * dd[4] * dd[8];
    } else {
      det = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
    }
    neg = neg || (det < 0.0);
    double wdj = TET ? ip_weight<TET>() * det : det;
    double det0 = det;
    if constexpr (AXI) {
      double r = 0.0;
#pragma unroll
      for (int a = 0; a < NN; a++) r += tabN<NN>(p, a) * X[a][0];
      det0 = det * r;
      wdj *= k2PI * r;
    }
    J->detJ0[e * NIP + p] = det0;
    if (J->detJ0f) J->detJ0f[e * NIP + p] = (float)det0;
#pragma unroll
    for (int a = 0; a < NN; a++) me[a] += J->rho0 * shapeN<NN, TET>(p, a) * wdj;
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
  const int W = J->slot_w;
  for (int k = 0; k < W; k++) m += J->elemF[(n * W + k) * DIM];
  if (J->ovf_ptr)
    for (int64_t o = J->N * W + J->ovf_ptr[n]; o < J->N * W + J->ovf_ptr[n + 1]; o++) m += J->elemF[o * DIM];
  J->mass[n] = m;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
void upload_tables(const double *hexN, const double *hexD, const double *quaN, const double *quaD) {
  upload_shape_tables(hexN, hexD, quaN, quaD);
  upload_tables_fp64(hexN, hexD, quaN, quaD);
}

template <int DIM, typename FT, int MAXB>
void launch_correct_b(JobDev *const *jobs, const Task *tasks, int nblk, int cs, int fold, hipStream_t s) {
  hipLaunchKernelGGL((k_correct<DIM, FT, MAXB>), dim3(nblk), dim3(kNodeBlock), 0, s, jobs, tasks, cs, fold);
}

template <int DIM>
void launch_correct_d(bool f32, int batch, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int fold,
                      hipStream_t s) {
  if (f32) {
    if (batch == 2)
      launch_correct_b<DIM, float, 2>(jobs, tasks, nblk, cs, fold, s);
    else if (batch == 4)
      launch_correct_b<DIM, float, 4>(jobs, tasks, nblk, cs, fold, s);
    else
      launch_correct_b<DIM, float, 8>(jobs, tasks, nblk, cs, fold, s);
  } else {
    if (batch == 2)
      launch_correct_b<DIM, double, 2>(jobs, tasks, nblk, cs, fold, s);
    else
      launch_correct_b<DIM, double, 4>(jobs, tasks, nblk, cs, fold, s);
  }
}

void launch_correct(int dim, bool f32, int batch, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int fold,
                    hipStream_t s) {
  if (nblk <= 0) return;
  if (dim == 3)
    launch_correct_d<3>(f32, batch, jobs, tasks, nblk, cs, fold, s);
  else
    launch_correct_d<2>(f32, batch, jobs, tasks, nblk, cs, fold, s);
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
  else if (type == DNLG_EL_TET4N_3D)
    hipLaunchKernelGGL((k_pi_elem<4, 3, false, true>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
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
  else if (type == DNLG_EL_TET4N_3D)
    hipLaunchKernelGGL((k_prep<4, 3, false, true>), dim3(nblk), dim3(256), 0, s, jobs, tasks);
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
