// energy.hip -- global energies of a job by a fixed-order compensated
// reduction on the device.
//
// Model::getTotalKineticEnergy (Model.C:612-625) sums mass * |v|^2 / 2 over
// the nodes (field0 speeds) in node order, one double accumulator.  The
// engine also reports the stored internal energy and the dissipated plastic
// work, sum_ip rho0 w detJ0 e_int and ... e_inel (x 2 pi for the
// axisymmetric quad, whose detJ0 carries the radius).  Here every term is
// added with Neumaier's compensation (a running sum and a running error
// term), block partials are merged with Knuth's TwoSum of the sums plus the
// errors, and the block count and the item-to-thread mapping depend only on
// the job's sizes: the result is the same bits on every run, and within a
// few ulp of the exactly rounded sum whatever the node count.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace dnlg {

// Scratch of dnlg_job_energies_static: the partials of at most
// kEnergyScratchBlocks blocks and the four results.  One copy per GPU, at
// its own address on each; host code uses the per-device address taken by
// energy_scratch_init, never this name (on the host it is only a stand-in).
static __device__ __align__(256) double scratch[6 * kEnergyScratchBlocks + 4];

namespace {

std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {nullptr};

struct Acc {
  double s, c;  // sum, accumulated rounding error
};

// Neumaier's step: s + x with the rounding error of that addition kept in c.
__device__ __forceinline__ void acc_add(Acc &a, double x) {
  const double t = a.s + x;
  a.c += (fabs(a.s) >= fabs(x)) ? (a.s - t) + x : (x - t) + a.s;
  a.s = t;
}

// Two accumulators: TwoSum of the sums (exact error), errors added.
__device__ __forceinline__ Acc acc_merge(const Acc &a, const Acc &b) {
  const double t = a.s + b.s;
  const double bp = t - a.s;
  const double err = (a.s - (t - bp)) + (b.s - bp);
  return Acc{t, a.c + b.c + err};
}

constexpr int kEnergyThreads = 256;
constexpr int kQ = 3;  // kinetic energy, internal energy, plastic work

// Fixed-shape tree over the block's threads; thread 0 ends with the total.
__device__ __forceinline__ void block_merge(Acc (&s)[kQ][kEnergyThreads], Acc (&a)[kQ]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kQ; q++) s[q][t] = a[q];
  __syncthreads();
  for (int w = kEnergyThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < kQ; q++) s[q][t] = acc_merge(s[q][t], s[q][t + w]);
    }
    __syncthreads();
  }
}

// Block b sums nodes [b pn, (b+1) pn) and integration points
// [b pg, (b+1) pg), thread t every kEnergyThreads-th item from its first;
// writes (sum, error) x 3 to part[6 b ..].  vol_scale: the Gauss weight
// (1, the tetrahedron's 1/6) times 2 pi for the axisymmetric quad.
template <int DIM>
__global__ __launch_bounds__(kEnergyThreads) void k_energy_partials(const JobDev *J, int64_t pn, int64_t pg,
                                                                   double vol_scale, double *part) {
  __shared__ Acc s[kQ][kEnergyThreads];
  const int64_t N = J->N, G = J->G;
  // between launch sequences the job's record is in control slot 0
  const double *v = J->fv[J->c[0].f0];
  Acc a[kQ] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  const int64_t n0 = (int64_t)blockIdx.x * pn, n1 = min(N, n0 + pn);
  for (int64_t n = n0 + threadIdx.x; n < n1; n += kEnergyThreads) {
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) v2 += v[d * N + n] * v[d * N + n];
    acc_add(a[0], (J->mass[n] * v2) / 2.0);
  }
  const double rho0 = J->rho0;
  const int64_t g0 = (int64_t)blockIdx.x * pg, g1 = min(G, g0 + pg);
  for (int64_t g = g0 + threadIdx.x; g < g1; g += kEnergyThreads) {
    const double vol = J->detJ0[g] * vol_scale;
    acc_add(a[1], rho0 * vol * J->eint[g]);
    acc_add(a[2], rho0 * vol * J->einel[g]);
  }
  block_merge(s, a);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kQ; q++) {
      part[6 * blockIdx.x + 2 * q] = s[q][0].s;
      part[6 * blockIdx.x + 2 * q + 1] = s[q][0].c;
    }
  }
}

// One block: partials in block order (thread t takes t, t + 256, ...), then
// the tree; out = {KE, IE, W_p, KE + IE}, each sum + error.
__global__ __launch_bounds__(kEnergyThreads) void k_energy_final(const double *part, int nb, double *out) {
  __shared__ Acc s[kQ][kEnergyThreads];
  Acc a[kQ] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  for (int b = threadIdx.x; b < nb; b += kEnergyThreads)
#pragma unroll
    for (int q = 0; q < kQ; q++) a[q] = acc_merge(a[q], Acc{part[6 * b + 2 * q], part[6 * b + 2 * q + 1]});
  block_merge(s, a);
  if (threadIdx.x == 0) {
    const double ke = s[0][0].s + s[0][0].c, ie = s[1][0].s + s[1][0].c, wp = s[2][0].s + s[2][0].c;
    const Acc tot = acc_merge(s[0][0], s[1][0]);
    out[0] = ke;
    out[1] = ie;
    out[2] = wp;
    out[3] = tot.s + tot.c;
  }
}

}  // namespace

int energy_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return DNLG_E_INVALID;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return DNLG_OK;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return DNLG_E_HIP;
  g_scratch_ptr[device] = (double *)ptr;
  return DNLG_OK;
}

double *energy_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

int energy_blocks(int64_t N, int64_t G) {
  const int64_t items = N > G ? N : G;
  const int64_t nb = (items + kEnergyItemsPerBlock - 1) / kEnergyItemsPerBlock;
  return (int)(nb < 1 ? 1 : nb);
}

void launch_energies(JobDev *job, int64_t N, int64_t G, int dim, double scale, double *part, double *out,
                     hipStream_t s) {
  const int nb = energy_blocks(N, G);
  const int64_t pn = (N + nb - 1) / nb, pg = (G + nb - 1) / nb;
  if (dim == 3)
    hipLaunchKernelGGL(k_energy_partials<3>, dim3(nb), dim3(kEnergyThreads), 0, s, job, pn, pg, scale, part);
  else
    hipLaunchKernelGGL(k_energy_partials<2>, dim3(nb), dim3(kEnergyThreads), 0, s, job, pn, pg, scale, part);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kEnergyThreads), 0, s, part, nb, out);
}

}  // namespace dnlg
