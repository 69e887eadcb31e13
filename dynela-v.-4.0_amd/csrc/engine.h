// engine.h -- device-side data layout of the explicit-dynamics engine.
//
// One job = one mesh.  Every job owns its device buffers (structure of
// arrays) and a device-resident JobDev record that also carries the solver
// scalars (time, time step, increment, ping-pong field index, status), so a
// cycle of the time loop needs no host round trip and can be captured in a
// HIP graph.  A batch is a list of JobDev pointers plus per-kernel block task
// lists, so one launch per phase advances any number of uneven jobs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dnlg.h"

namespace dnlg {

// Constants of Sources/dnlKernel/Macros.h:37-56.
constexpr double kPI = 3.14159265358979324;
constexpr double k2PI = 6.28318530717958648;
constexpr double kSqrt3 = 1.73205080756887729;
constexpr double kSqrt23 = 0.816496580927726034;
constexpr double kSqrt32 = 1.22474487139158894;
constexpr double kEps = 2.2204460492503131e-16;

// Reference NR settings (Element.C:407-430).
constexpr double kGammaInitial = 1e-8;
constexpr double kTolNR = 1e-8;
constexpr int kItMax = 250;

__host__ __device__ __forceinline__ double sqd(double a) { return a * a; }

// Element types: spatial dimensions, nodes and integration points.
__host__ __device__ __forceinline__ bool type_known(int type) {
  return type == DNLG_EL_HEX8N_3D || type == DNLG_EL_QUA4N_AX || type == DNLG_EL_QUA4N_2D || type == DNLG_EL_TET4N_3D;
}
__host__ __device__ __forceinline__ int type_dim(int type) {
  return (type == DNLG_EL_HEX8N_3D || type == DNLG_EL_TET4N_3D) ? 3 : 2;
}
__host__ __device__ __forceinline__ int type_nodes(int type) { return type == DNLG_EL_HEX8N_3D ? 8 : 4; }
__host__ __device__ __forceinline__ int type_ips(int type) {
  return type == DNLG_EL_HEX8N_3D ? 8 : (type == DNLG_EL_TET4N_3D ? 1 : 4);
}

// Factor from an integration point's reference det J to its volume: the
// Gauss weight (1, the tetrahedron's 1/6) times 2 pi for the axisymmetric
// quad, whose detJ0 carries the radius.
__host__ __device__ __forceinline__ double type_vol_scale(int type) {
  return type == DNLG_EL_QUA4N_AX ? k2PI : (type == DNLG_EL_TET4N_3D ? 1.0 / 6 : 1.0);
}

// Time-step methods that run the power iteration (fully or under-integrated
// element stiffness).
__host__ __device__ __forceinline__ bool is_pi(int dt_method) {
  return dt_method == DNLG_DT_POWER_ITERATION || dt_method == DNLG_DT_POWER_ITERATION_UNDER;
}

// Precisions with FP32 element state (DNLG_PREC_MIXED, DNLG_PREC_MIXED_RR32).
__host__ __device__ __forceinline__ bool is_mixed(int prec) { return prec != DNLG_PREC_FP64; }

struct Boundary {
  int32_t kind;  // DNLG_BC_*
  int32_t ramp;  // DNLG_RAMP_*
  double v[3];
  double rmin, rmax;
};

// Solver scalars of one cycle of a job (Solver.h:45-50 plus engine flags).
// A job keeps two records: while a launch sequence runs, the host names for
// each launch the slot that holds its cycle's record, and the kernel that
// opens a cycle writes the other slot (so no launch rewrites a record that
// its own blocks read).  Between sequences the record is in slot 0.
struct Control {
  double t;      // model->currentTime
  double dt;     // Solver::timeStep (persists, clipped at the end)
  double up_to;  // target of the current advance
  int64_t inc;   // Solver::currentIncrement
  int32_t f0;    // which ping-pong slot is the reference's field0
  int32_t running;      // this cycle advances the job
  int32_t courant_due;  // Courant recompute after this cycle
  int32_t density_due;  // element kernel writes IP densities this cycle
  int32_t preloop;      // element pass of Explicit.C:158-178
  int32_t pending;      // the last cycle ran and awaits its end-of-step
};

// Error codes (DNLG_E_*, sticky, the first one wins), by the parity of the
// increment of the cycle that raised them: a launch that opens cycle i + 1
// reads cycle i's word, which no launch of cycle i + 1 writes.  A job stops
// after its first failing cycle, so at most one word is ever set.  On a
// cache line of its own: written by device atomics while other blocks store
// control records.
struct alignas(128) Sync {
  int32_t st[2];
};

// Power-iteration time step (Model::computePowerIterationTimeStep,
// Model.C:942-1009) of one job: the iterate lives in pi_ev[cur], normalised
// as pi_ev[cur] / scale (the reference divides in place by the same value).
struct PiState {
  double fmax;    // last maximum (Model.h:46), 0 before the first call
  double scale;   // divisor of the stored iterate (1 for the starting vector)
  int32_t cur;    // which pi_ev buffer holds the iterate
  int32_t iter;   // iterations of the current call
  int32_t call;   // a time-step computation is in progress
  int32_t done;   // ... and has converged (or failed)
  int32_t stall;  // a cycle ended before its iteration converged
  int32_t pad;
};

// Result (VTK) files of DynELA::writeResultFile (DynELA.C:489-503), called
// from Explicit::endStep: after a cycle whose new time t has
// next <= t <= end a file is written and next advances one increment.  On
// the device the cycle's commit sets `hold`; a held job opens no further
// cycle until the host has written the file from its state (the cycle is
// closed by then: field0 of the written cycle is the other ping-pong slot)
// and cleared the flag.
struct ResultSched {
  double next, end, inc;
  int32_t on, hold;
};

// History files (HistoryFile::save, HistoryFile.C:283-315).  A file takes a
// row at time t unless start < t < next; taking one advances next by one
// save interval.  Rows go to a device ring of `cap` rows (row r at r % cap,
// 1 + n_items doubles: time, then the items in file order) that the host
// drains between launch sequences; next and rows change only in
// k_hist_commit.
struct HistFile {
  double start, next, save;
  int64_t rows;       // rows taken so far
  double *ring;
  int32_t cap, n_items;
  int32_t ke0, n_ke;  // kinetic-energy columns: hke_col[ke0 .. ke0 + n_ke)
};

// Item evaluations (HistoryFileNodeItem / ElementItem / GlobalItem::getValue,
// HistoryFile.C:46-99).
enum HistOp : int32_t {
  kHZero = 0,    // a field the reference's getValue does not know: 0
  kHNodal = 1,   // Node::fieldScalar nodal part (Node.C:311-330), src HistNodal
  kHNodeIP = 2,  // IP values extrapolated to the node (Node.C:244-300, 350-364)
  kHElemIP = 3,  // Element::getIntPointValue (Element.C:1195-1238)
  kHDt = 4,      // Solver::timeStep (realTimeStep, timeStep)
  kHKe = 5,      // Model::getTotalKineticEnergy, written by k_hist_commit
};
enum HistNodal : int32_t { kHCoords = 0, kHDisp = 1, kHDispInc = 2, kHSpeed = 3, kHSpeedInc = 4, kHMass = 5 };
enum HistIP : int32_t {
  kHEp = 0, kHDep = 1, kHGam = 2, kHGamc = 3, kHSy = 4, kHT = 5, kHRho = 6, kHPres = 7,
  kHEint = 8, kHVm = 9, kHSig = 10, kHEps = 11, kHEpsP = 12,
};
constexpr int kHNorm = 6;  // comp of a tensor item: its norm (vectors: 3)

// Blocks per job of the history evaluation launch: items and kinetic-energy
// node ranges are split over at most this many blocks each (grid-stride
// loops inside), so a cycle that takes no row costs a small launch.
constexpr int kHistMaxBlocks = 256;
__host__ __device__ inline int hist_blocks(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > kHistMaxBlocks ? kHistMaxBlocks : b));
}

struct HistItem {
  int32_t file, col;  // owning file and column in its row (1 + position)
  int32_t op, src;    // HistOp, HistNodal / HistIP
  int32_t comp;       // vector 0..2 or 3 (norm); tensor [xx,xy,xz,yy,yz,zz] or 6
  int32_t ip;         // element items: IP index, -1 = mean over the IPs
  int64_t index;      // node or element
  int32_t r0, r1;     // node-from-IP items: incidence pairs hinc[r0 .. r1)
};

struct JobDev {
  Control c[2];
  Sync sync;
  PiState pi;
  ResultSched res;
  // sizes
  int32_t type, dim, nn, nip;
  int64_t N, NE, G;  // nodes, elements, integration points
  int32_t prec, track, law, dt_freq, dt_method, stress_method;
  int32_t nr_max_it;  // Newton iterations of the radial return (250 by default)
  // material (Material.h:110-169) and law parameters
  double rho0, young, nu, twoG, K, heatFrac, eta, cp, T0;
  double lp[8];
  const double *ltab;  // DNLG_LAW_TABULAR: strains [0, n_ltab), yield stresses [n_ltab, 2 n_ltab)
  int32_t n_ltab;
  // FP32 copies for the mixed path's radial return, and 1 / rho0
  float lpf[8];
  float inv_dTf;  // 1 / (Tm - T0) of Johnson-Cook
  double inv_rho0;
  // Chung-Hulbert (Explicit.C:86-104) and time-step settings
  double alphaM, beta, gammaCH, omegaS, safety;
  // nodal state, [dim][N]
  double *x, *x0, *mass;  // current and initial coordinates, lumped mass
  double *fu[2], *fv[2], *fa[2];
  const int32_t *bc_ptr;  // [N+1] constant-BC list per node
  const int32_t *bc_idx;
  const Boundary *bnd;
  const int32_t *bcn;     // nodes that carry constant BCs, ascending
  int32_t n_bcn;
  int32_t bc_flags;       // 1: a constant speed BC, 2: a ramped one
  // Incidence slots, node-major in element order: node n's first slot_w
  // incidences are slots n slot_w .. n slot_w + slot_w - 1 (slots past its
  // incidence count stay zero), further ones (a node with more than slot_w
  // incident elements) slots N slot_w + ovf_ptr[n] .. N slot_w + ovf_ptr[n+1]
  // - 1; ovf_ptr is null when no node has more.
  int32_t slot_w;
  const int32_t *ovf_ptr;  // [N+1] or null
  const int32_t *e2slot;   // [NE*nn] incidence slot of (element, local node)
  // elements
  const int32_t *conn;  // [NE][nn]
  double *elemF;        // [slot][dim] element internal forces (per incidence slot)
  float *elemF32;       // mixed-precision variant of elemF
  double *cmin;         // Courant partial minima, one per Courant block
  int32_t n_cmin;
  // power iteration (DNLG_DT_POWER_ITERATION[_UNDER] only): iterates [dim][N] x 2,
  // per-node-block partial maxima
  double *pi_ev[2];
  double *pi_bmax;
  int32_t n_pibmax;
  // integration-point state, index e*nip + p; tensors as [6][G]
  double *sig;  // FP64 stress (FP64 mode)
  float *sig32; // FP32 stress (mixed mode)
  double *ep, *eint, *einel, *T;  // FP64 in both modes
  double *dep, *sy, *gam, *gamc;  // FP64 mode
  float *dep32, *sy32, *gam32;  // mixed mode (its cumulative gamma is ep / sqrt(2/3))
  double *rho, *detJ0;
  float *detJ0f;
  double *eps, *epsP;  // optional tensors (track_tensors)
  // history files (none: n_hf == 0)
  HistFile *hf;
  const HistItem *hi;
  const int2 *hinc;       // (element, local node) pairs of node-from-IP items
  const int32_t *hke_col;  // kinetic-energy columns, per file (HistFile::ke0)
  double *hke;            // kinetic-energy partial sums, hist_blocks(N) of them
  int32_t n_hf, n_hi;
  int32_t density_always;  // IP densities every cycle (a history item reads them)
  // node -> (element, local node) lists in element order (Node::elements),
  // for IP fields extrapolated to every node; built on first use
  const int32_t *nf_ptr;
  const int2 *nf_pairs;
};

// Address spaces for the kernels.  A job record is read-only while a kernel
// runs (apart from the sticky status word, which only the atomic setter
// touches), and uniform over a block: reading it through a constant-space
// pointer lets the compiler use scalar loads, which do not wait behind the
// vector loads of the element / node data.  Data arrays are accessed through
// global-space pointers (global instead of flat instructions).
#define DNLG_GLOBAL __attribute__((address_space(1)))
#define DNLG_CONST __attribute__((address_space(4)))
typedef const DNLG_CONST JobDev CJobDev;
template <typename T>
__device__ __forceinline__ DNLG_GLOBAL T *gbl(T *p) {
  return (DNLG_GLOBAL T *)p;
}

// Reference at a uniform base plus a 32-bit byte offset: the global
// instructions' scalar-base + 32-bit vector-offset form, so that a lane
// forms no 64-bit address of its own.  Kernels give the base the block's
// first item, which keeps the offsets of per-item arrays lane-local.
// The base passes through readfirstlane (it is uniform), which keeps the
// compiler from folding uniform and per-lane terms into one 64-bit
// per-lane address.
template <typename T>
__device__ __forceinline__ DNLG_GLOBAL T &at(const DNLG_GLOBAL T *base, uint32_t byte_off) {
  const uint64_t b = (uint64_t)base;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b), hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return *(DNLG_GLOBAL T *)((DNLG_GLOBAL char *)(((uint64_t)hi << 32) | lo) + byte_off);
}

// Threads per block of the node kernels (corrector, mass assembly): one node
// per thread, contiguous node ranges per block.
constexpr int kNodeBlock = 256;

// One block of work: job index into the batch, local block index within the
// job, the half-open item range [first, end), and the job's record (filled
// in when the task list is uploaded, so a block reaches its job with one
// load less).
struct Task {
  int32_t job;
  int32_t blk;
  int64_t first;
  int64_t end;
  JobDev *jd = nullptr;
};

// The block's task, read through the constant space.
__device__ __forceinline__ Task load_task(const Task *tasks) {
  const DNLG_CONST Task *t = (const DNLG_CONST Task *)tasks + blockIdx.x;
  Task tk;
  tk.job = t->job;
  tk.blk = t->blk;
  tk.first = t->first;
  tk.end = t->end;
  tk.jd = t->jd;
  return tk;
}

}  // namespace dnlg
