// history.hip -- history-file rows taken on the device.
//
// The reference writes its history files from Explicit::endStep, after the
// time update and before the field swap (Explicit.C:539-552 calling
// Model::writeHistoryFiles, Model.C:917-924), and once after the pre-loop
// pass (Explicit.C:182).  Here, for the jobs of a batch that have history
// files, two launches follow the corrector of every cycle (and the
// pre-loop pass):
//   k_hist_eval    the item values, into the rows of the files that take a
//                  row at t_{n+1}, and kinetic-energy partial sums when such
//                  a file has a kineticEnergy item (at most kHistMaxBlocks
//                  blocks per job for each)
//   k_hist_commit  one block per job: time column, kinetic energy, next save
//                  time and row count of the files that take a row; the
//                  result-file (VTK) schedule, which holds the job when a
//                  file is due (DynELA::writeResultFile, DynELA.C:489-503)
// At that point the cycle's field0 (the reference's field0 before the swap)
// is fu/fv/fa[c.f0], coordinates and IP values are the new ones.  The host
// drains the rings between launch sequences and writes the files.
#include <hip/hip_runtime.h>

#include <cmath>

#include "control.h"
#include "kernels.h"

namespace dnlg {

// IP-to-node extrapolation (the integrationPointsToNode vectors of
// ElQua4NAx.C:30-60 and ElHex8N3D.C:35-90): the inverse of the shape
// function matrix N[ip][node] at the Gauss points.  Along one direction the
// two Gauss values g = +-1/sqrt3 extrapolate to the element end with weight
// (1 + sqrt3) / 2 for the IP on the end's side and (1 - sqrt3) / 2 for the
// other one; the product over the directions gives a^k b^(dim-k), k = the
// directions in which IP and node lie on the same side.  IP p sits at the
// node sign pattern of node p scaled by 1/sqrt3.
__constant__ double c_hexE[8][8];  // [node][ip]
__constant__ double c_quaE[4][4];

void extrapolation_tables(double hexE[64], double quaE[16]) {
  const double a = (1.0 + std::sqrt(3.0)) / 2, b = (1.0 - std::sqrt(3.0)) / 2;
  const int hs[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int n = 0; n < 8; n++)
    for (int p = 0; p < 8; p++) {
      double w = 1.0;
      for (int c = 0; c < 3; c++) w *= hs[n][c] == hs[p][c] ? a : b;
      hexE[n * 8 + p] = w;
    }
  const int qs[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  for (int n = 0; n < 4; n++)
    for (int p = 0; p < 4; p++) {
      double w = 1.0;
      for (int c = 0; c < 2; c++) w *= qs[n][c] == qs[p][c] ? a : b;
      quaE[n * 4 + p] = w;
    }
}

void upload_extrapolation() {
  double hexE[64], quaE[16];
  extrapolation_tables(hexE, quaE);
  hipMemcpyToSymbol(HIP_SYMBOL(c_hexE), hexE, sizeof(hexE));
  hipMemcpyToSymbol(HIP_SYMBOL(c_quaE), quaE, sizeof(quaE));
}

namespace {

// nip == 1: the tetrahedron (ElTet4N3D.C:33-60), whose nodes' vectors hold
// four entries while it has one integration point: the first entry,
// (1 + 3 sqrt5)/4 at node 0 and (1 - sqrt5)/4 at the others, is the one used.
__device__ __forceinline__ double extrap(int nip, int node, int p) {
  if (nip == 1) return node == 0 ? 1.9270509831248424 : -0.30901699437494745;
  return nip == 8 ? c_hexE[node][p] : c_quaE[node][p];
}

// HistoryFile::save (HistoryFile.C:294-296): a row unless start < t < next.
__device__ __forceinline__ bool hist_due(const DNLG_GLOBAL HistFile &f, double t) {
  return !(t > f.start && t < f.next);
}

// The moment a row would be taken at: the pre-loop (t, its forced time
// step), or the end of the cycle in slot cs (t + dt), which must have run
// without error (the reference stops before its endStep otherwise).
__device__ __forceinline__ bool hist_moment(CJobDev *C, int cs, int preloop, double &t, double &dt, int &f0) {
  const Control c = load_control(C, cs);
  if (preloop) {
    if (job_status(C) != 0) return false;
    t = c.t;
  } else {
    const int st = (c.inc & 1) ? C->sync.st[1] : C->sync.st[0];
    if (!c.running || st != 0) return false;
    t = c.t + c.dt;
  }
  dt = c.dt;
  f0 = c.f0;
  return true;
}

__device__ __forceinline__ void load_tensor(CJobDev *C, int src, int64_t g, double s[6]) {
  const int64_t G = C->G;
  if (src == kHEps || src == kHEpsP) {
    const DNLG_GLOBAL double *t = gbl(src == kHEps ? C->eps : C->epsP);
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = t[k * G + g];
  } else if (is_mixed(C->prec)) {
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = (double)gbl(C->sig32)[k * G + g];
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = gbl(C->sig)[k * G + g];
  }
}

// SymTensor2::norm (SymTensor2.h:362-367)
__device__ __forceinline__ double sym_norm(const double s[6]) {
  return sqrt(sqd(s[0]) + sqd(s[3]) + sqd(s[5]) + 2.0 * (sqd(s[1]) + sqd(s[2]) + sqd(s[4])));
}

// One IP's value of a field (Element::getIntPointValueExtract,
// Element.C:1195-1221; the mixed path's FP32 arrays widened).
__device__ double ip_value(CJobDev *C, int src, int comp, int64_t g) {
  const bool mixed = is_mixed(C->prec);
  switch (src) {
    case kHEp: return gbl(C->ep)[g];
    case kHDep: return mixed ? (double)gbl(C->dep32)[g] : gbl(C->dep)[g];
    case kHGam: return mixed ? (double)gbl(C->gam32)[g] : gbl(C->gam)[g];
    // the mixed path keeps the plastic strain only: gammaCumulate = ep / sqrt(2/3)
    case kHGamc: return mixed ? gbl(C->ep)[g] / kSqrt23 : gbl(C->gamc)[g];
    case kHSy: return mixed ? (double)gbl(C->sy32)[g] : gbl(C->sy)[g];
    case kHT: return gbl(C->T)[g];
    case kHRho: return gbl(C->rho)[g];
    case kHEint: return gbl(C->eint)[g];
    default: break;
  }
  double s[6];
  load_tensor(C, src, g, s);
  if (src == kHPres) return (s[0] + s[3] + s[5]) / 3.0;
  if (src == kHVm)  // dnlVonMises (Macros.h:71-72)
    return sqrt((sqd(s[0] - s[3]) + sqd(s[0] - s[5]) + sqd(s[3] - s[5]) + 6 * (sqd(s[1]) + sqd(s[2]) + sqd(s[4]))) / 2);
  return comp == kHNorm ? sym_norm(s) : s[comp];
}

__device__ double item_value(CJobDev *C, const HistItem &it, int f0, double dt) {
  const int64_t N = C->N;
  const int nip = C->nip;
  switch (it.op) {
    case kHNodal: {
      const int64_t n = it.index;
      if (it.src == kHMass) return gbl(C->mass)[n];
      double v[3] = {0.0, 0.0, 0.0};
      const DNLG_GLOBAL double *a = nullptr;
      switch (it.src) {
        case kHCoords:
        case kHDisp: a = gbl(C->x); break;
        case kHDispInc: a = gbl(C->fu[f0]); break;
        case kHSpeed: a = gbl(C->fv[f0]); break;
        default: a = gbl(C->fa[f0]); break;
      }
      for (int d = 0; d < C->dim; d++) {
        v[d] = a[d * N + n];
        // the node displacement sum of Explicit.C:530-531 is x - x0
        if (it.src == kHDisp) v[d] -= gbl(C->x0)[d * N + n];
      }
      // Vec3D::norm (Vec3D.h:400-404)
      return it.comp == 3 ? sqrt(sqd(v[0]) + sqd(v[1]) + sqd(v[2])) : v[it.comp];
    }
    case kHNodeIP: {
      // sum over the node's elements and their IPs of the extrapolation
      // weight times the IP value, over the element count
      const bool tnorm = it.comp == kHNorm && (it.src == kHSig || it.src == kHEps || it.src == kHEpsP);
      double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int r = it.r0; r < it.r1; r++) {
        const int2 pr = C->hinc[r];
        for (int p = 0; p < nip; p++) {
          const double w = extrap(nip, pr.y, p);
          const int64_t g = (int64_t)pr.x * nip + p;
          if (tnorm) {
            double s[6];
            load_tensor(C, it.src, g, s);
#pragma unroll
            for (int k = 0; k < 6; k++) acc[k] += w * s[k];
          } else {
            acc[0] += w * ip_value(C, it.src, it.comp, g);
          }
        }
      }
      const double cnt = (double)(it.r1 - it.r0);
      if (!tnorm) return acc[0] / cnt;
#pragma unroll
      for (int k = 0; k < 6; k++) acc[k] = acc[k] / cnt;
      return sym_norm(acc);
    }
    case kHElemIP: {
      const int64_t e = it.index;
      if (it.ip >= 0) return ip_value(C, it.src, it.comp, e * nip + it.ip);
      double v = 0.0;
      for (int p = 0; p < nip; p++) v += ip_value(C, it.src, it.comp, e * nip + p);
      return v / nip;
    }
    case kHDt: return dt;
    default: return 0.0;
  }
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x / 64, nw = blockDim.x / 64;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  T s = 0;
  for (int k = 0; k < nw; k++) s += lds[k];
  return s;
}

}  // namespace

// Items (task blk >= 0: a range of the job's items) and kinetic-energy
// partial sums (blk < 0: node range -blk - 1 of the job's hist_blocks(N)),
// both as grid-stride loops over the task's range.  Model::
// getTotalKineticEnergy (Model.C:612-625) on the cycle's field0 speeds:
// m (v . v) / 2 per node.
__global__ __launch_bounds__(256) void k_hist_eval(JobDev *const *jobs, const Task *tasks, int cs, int preloop) {
  __shared__ double lds[256 / 64];
  const Task tk = load_task(tasks);
  CJobDev *C = (CJobDev *)tk.jd;
  double t, dt;
  int f0;
  if (!hist_moment(C, cs, preloop, t, dt, f0)) return;
  if (tk.blk >= 0) {
    for (int64_t k = tk.first + threadIdx.x; k < tk.end; k += blockDim.x) {
      const HistItem it = C->hi[k];
      if (it.op == kHKe) continue;
      const DNLG_GLOBAL HistFile &f = gbl(C->hf)[it.file];
      if (!hist_due(f, t)) continue;
      const double v = item_value(C, it, f0, dt);
      gbl(f.ring)[(f.rows % f.cap) * (1 + f.n_items) + it.col] = v;
    }
    return;
  }
  bool any = false;
  for (int i = 0; i < C->n_hf; i++) {
    const DNLG_GLOBAL HistFile &f = gbl(C->hf)[i];
    any = any || (f.n_ke > 0 && hist_due(f, t));
  }
  if (!any) return;
  const int64_t N = C->N;
  const DNLG_GLOBAL double *v = gbl(C->fv[f0]);
  double e = 0.0;
  for (int64_t n = tk.first + threadIdx.x; n < tk.end; n += blockDim.x) {
    double v2 = 0.0;
    for (int d = 0; d < C->dim; d++) v2 += sqd(v[d * N + n]);
    e += (gbl(C->mass)[n] * v2) / 2.0;
  }
  const double s = block_sum(e, lds);
  if (threadIdx.x == 0) gbl(C->hke)[-tk.blk - 1] = s;
}

__global__ __launch_bounds__(64) void k_hist_commit(JobDev *const *jobs, const Task *tasks, int cs, int preloop) {
  __shared__ double s_ke;
  const Task tk = load_task(tasks);
  JobDev *J = tk.jd;
  CJobDev *C = (CJobDev *)J;
  double t, dt;
  int f0;
  if (!hist_moment(C, cs, preloop, t, dt, f0)) return;
  if (threadIdx.x == 0) {
    bool any = false;
    for (int i = 0; i < C->n_hf; i++) {
      const DNLG_GLOBAL HistFile &f = gbl(C->hf)[i];
      any = any || (f.n_ke > 0 && hist_due(f, t));
    }
    double ke = 0.0;
    if (any) {
      const int nb = hist_blocks(C->N);
      for (int b = 0; b < nb; b++) ke += gbl(C->hke)[b];
    }
    s_ke = ke;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < C->n_hf; i += blockDim.x) {
    DNLG_GLOBAL HistFile &f = gbl(J->hf)[i];
    if (!hist_due(f, t)) continue;
    DNLG_GLOBAL double *row = gbl(f.ring) + (f.rows % f.cap) * (1 + f.n_items);
    row[0] = t;
    for (int k = 0; k < f.n_ke; k++) row[gbl(C->hke_col)[f.ke0 + k]] = s_ke;
    // HistoryFile.C:298-299: one interval further, however far t has gone
    f.next += f.save;
    f.rows += 1;
  }
  // DynELA::writeResultFile: after the history rows, in the cycle's endStep
  if (!preloop && threadIdx.x == 0 && C->res.on && t >= C->res.next && t <= C->res.end) {
    DNLG_GLOBAL ResultSched &r = *gbl(&J->res);
    r.hold = 1;
    r.next = C->res.next + C->res.inc;
  }
}

// Node::fieldScalar / fieldSymTensor2 of an IP field at every node
// (Node.C:244-300, 350-364, 411-425): the node's elements in element order,
// per element the extrapolation weights of the node's local position times
// the IP values, over the element count.
__global__ __launch_bounds__(256) void k_node_field(JobDev *J, int src, int comp, double *out) {
  CJobDev *C = (CJobDev *)J;
  const int64_t N = C->N;
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  const int nip = C->nip;
  const int r0 = C->nf_ptr[n], r1 = C->nf_ptr[n + 1];
  const bool tensor = comp < 0 || (comp == kHNorm && (src == kHSig || src == kHEps || src == kHEpsP));
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int r = r0; r < r1; r++) {
    const int2 pr = C->nf_pairs[r];
    for (int p = 0; p < nip; p++) {
      const double w = extrap(nip, pr.y, p);
      const int64_t g = (int64_t)pr.x * nip + p;
      if (tensor) {
        double s[6];
        load_tensor(C, src, g, s);
#pragma unroll
        for (int k = 0; k < 6; k++) acc[k] += w * s[k];
      } else {
        acc[0] += w * ip_value(C, src, comp, g);
      }
    }
  }
  const double cnt = (double)(r1 - r0);
  if (!tensor) {
    gbl(out)[n] = acc[0] / cnt;
    return;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) acc[k] = acc[k] / cnt;
  if (comp < 0) {
#pragma unroll
    for (int k = 0; k < 6; k++) gbl(out)[k * N + n] = acc[k];
  } else {
    gbl(out)[n] = sym_norm(acc);
  }
}

void launch_node_field(JobDev *job, int64_t N, int src, int comp, double *out, hipStream_t s) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_node_field, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, job, src, comp, out);
}

void launch_history(JobDev *const *jobs, const Task *eval, int n_eval, const Task *per_job, int n_jobs, int cs,
                    int preloop, hipStream_t s) {
  if (n_jobs <= 0) return;
  if (n_eval > 0) hipLaunchKernelGGL(k_hist_eval, dim3(n_eval), dim3(256), 0, s, jobs, eval, cs, preloop);
  hipLaunchKernelGGL(k_hist_commit, dim3(n_jobs), dim3(64), 0, s, jobs, per_job, cs, preloop);
}

}  // namespace dnlg
