// api.hip -- C ABI of libdnlg.so: contexts, jobs, batched solve, read-back.
//
// Host C++ replaces the reference's Model/Explicit objects for the explicit
// solve: Model::initSolve (Model.C:360-521) becomes dnlg_job_create,
// Solver::solve / Explicit::solve (Solver.h:70, Explicit.C:126-253) becomes
// dnlg_job_solve / dnlg_batch_solve.  All per-cycle work runs on one HIP
// stream per context; the host only checks progress between chunks of
// cycles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dnlg.h"
#include "batch.h"
#include "kernels.h"
#include "textfmt.h"

using namespace dnlg;

namespace {

int g_tables_uploaded[64] = {0};
std::mutex g_tables_mu;

// Gauss-point shape tables: closed forms of ElQua4NAx.C:63-110 and
// ElHex8N3D.C:93-231 (products of (1 +- 1/sqrt3) factors).
void make_tables(double hexN[64], double hexD[192], double quaN[16], double quaD[32]) {
  const int hs[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                        {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  const double hN[4] = {(9.0 - 5.0 * std::sqrt(3.0)) / 36, (3 - std::sqrt(3.0)) / 36, (3 + std::sqrt(3.0)) / 36,
                        (9.0 + 5.0 * std::sqrt(3.0)) / 36};
  const double hD[3] = {(2.0 - std::sqrt(3.0)) / 12, 1.0 / 12, (2.0 + std::sqrt(3.0)) / 12};
  for (int p = 0; p < 8; p++)
    for (int a = 0; a < 8; a++) {
      int k = 0;
      for (int c = 0; c < 3; c++) k += (hs[a][c] == hs[p][c]);
      hexN[p * 8 + a] = hN[k];
      for (int c = 0; c < 3; c++) {
        int kk = 0;
        for (int o = 0; o < 3; o++)
          if (o != c) kk += (hs[a][o] == hs[p][o]);
        hexD[(p * 8 + a) * 3 + c] = hs[a][c] > 0 ? hD[kk] : -hD[kk];
      }
    }
  const int qs[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  const double qN[3] = {(2.0 - std::sqrt(3.0)) / 6, 1.0 / 6, (2.0 + std::sqrt(3.0)) / 6};
  const double qD[2] = {(3.0 - std::sqrt(3.0)) / 12, (3.0 + std::sqrt(3.0)) / 12};
  for (int p = 0; p < 4; p++)
    for (int a = 0; a < 4; a++) {
      int k = 0;
      for (int c = 0; c < 2; c++) k += (qs[a][c] == qs[p][c]);
      quaN[p * 4 + a] = qN[k];
      for (int c = 0; c < 2; c++) {
        const int o = 1 - c;
        const double v = qD[qs[a][o] == qs[p][o] ? 1 : 0];
        quaD[(p * 4 + a) * 2 + c] = qs[a][c] > 0 ? v : -v;
      }
    }
}

int fail(dnlg_ctx *ctx, int st, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return st;
}

#define HIPCHK(ctx, expr)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      return fail((ctx), e_ == hipErrorOutOfMemory ? DNLG_E_OOM : DNLG_E_HIP,              \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    }                                                                                      \
  } while (0)

template <typename T>
int dalloc(dnlg_job *job, T **p, int64_t count, const T *init = nullptr) {
  *p = nullptr;
  if (count <= 0) return DNLG_OK;
  void *ptr = nullptr;
  HIPCHK(job->ctx, hipMalloc(&ptr, sizeof(T) * count));
  job->allocs.push_back(ptr);
  *p = (T *)ptr;
  if (init) {
    HIPCHK(job->ctx, hipMemcpy(ptr, init, sizeof(T) * count, hipMemcpyHostToDevice));
  } else {
    HIPCHK(job->ctx, hipMemset(ptr, 0, sizeof(T) * count));
  }
  return DNLG_OK;
}

// Incidence slots of a mesh (the layout JobDev describes): slot_w = the
// largest incidence count, at most nn (the count at an interior node of a
// structured mesh); node n's k-th incident (element, local node) in element
// order goes to slot n slot_w + k for k < slot_w, else to the overflow slot
// N slot_w + ovf[n] + k - slot_w.  Fills e2slot (and ovf, left empty when
// no node has more than slot_w incidences) when given; returns the number
// of slots.
int64_t incidence_slots(const int32_t *conn, int64_t N, int64_t NE, int nn, int32_t *slot_w,
                        std::vector<int32_t> *ovf, std::vector<int32_t> *e2slot) {
  std::vector<int32_t> cnt(N, 0);
  for (int64_t k = 0; k < NE * nn; k++) cnt[conn[k]]++;
  int32_t mx = 1;
  for (int64_t n = 0; n < N; n++) mx = std::max(mx, cnt[n]);
  const int32_t W = std::min(mx, (int32_t)nn);
  std::vector<int32_t> optr(N + 1, 0);
  for (int64_t n = 0; n < N; n++) optr[n + 1] = optr[n] + std::max(cnt[n] - W, 0);
  const int64_t total = N * W + optr[N];
  if (slot_w) *slot_w = W;
  if (e2slot) {
    e2slot->assign(NE * nn, 0);
    std::vector<int32_t> fill(N, 0);
    for (int64_t e = 0; e < NE; e++)
      for (int a = 0; a < nn; a++) {
        const int32_t n = conn[e * nn + a], k = fill[n]++;
        (*e2slot)[e * nn + a] = k < W ? (int32_t)(n * W + k) : (int32_t)(N * W + optr[n] + k - W);
      }
  }
  if (ovf) {
    ovf->clear();
    if (optr[N] > 0) *ovf = std::move(optr);
  }
  return total;
}

// Device bytes of a job.  n_slots < 0: the incidence slots are not known
// (no connectivity at hand) and are bounded by NE nn + N nn.
int64_t footprint(int type, int64_t N, int64_t NE, int prec, int track, int64_t n_slots = -1) {
  const int dim = type_dim(type);
  const int nn = type_nodes(type);
  const int64_t G = NE * type_ips(type);
  if (n_slots < 0) n_slots = NE * nn + N * nn;
  int64_t b = 0;
  b += N * dim * 8 * 8;            // x, x0, fu/fv/fa x2
  b += N * 8;                      // mass
  b += (N + 1) * 4 * 2 + NE * nn * 4 * 2;  // overflow ranges, BC ranges, connectivity, slots
  b += N * 4;                      // BC node list (upper bound)
  b += n_slots * dim * 8;          // elemF
  if (is_mixed(prec)) {
    b += n_slots * dim * 4;        // elemF32
    b += G * (6 * 4 + 3 * 4 + 4 * 8 + 8 + 8 + 4);
  } else {
    b += G * (6 * 8 + 8 * 8 + 8 + 8);
  }
  if (track) b += G * 12 * 8;
  b += N * dim * 8 * 2 + ((N + 255) / 256) * 8;  // power-iteration iterates (upper bound: any method)
  b += ((NE + 255) / 256) * 8 + sizeof(JobDev) + 4096;
  return b;
}


// Labels of Field::FieldLabel (Field.C vtkNames): the enum identifiers,
// yieldStress labelled "yield".
std::string vtk_label(int f) {
  static const char *vec[5] = {"coords", "disp", "dispInc", "speed", "speedInc"};
  static const char *sca[15] = {"density", "gamma", "gammaCumulate", "internalEnergy", "inelasticEnergy",
                                "kineticEnergy", "mass", "plasticStrain", "plasticStrainRate", "pressure",
                                "realTimeStep", "T", "timeStep", "vonMises", "yield"};
  static const char *ten[5] = {"PlasticStrain", "PlasticStrainInc", "Strain", "StrainInc", "Stress"};
  static const char *vc[4] = {"", "X", "Y", "Z"};
  static const char *tc[10] = {"", "XX", "XY", "XZ", "YX", "YY", "YZ", "ZX", "ZY", "ZZ"};
  if (f < 0 || f >= DNLG_FIELD_COUNT) return "";
  if (f < DNLG_FIELD_DENSITY) return std::string(vec[f / 4]) + vc[f % 4];
  if (f < DNLG_FIELD_PLASTIC_STRAIN_TENSOR) return sca[f - DNLG_FIELD_DENSITY];
  const int k = f - DNLG_FIELD_PLASTIC_STRAIN_TENSOR;
  return std::string(ten[k / 10]) + tc[k % 10];
}

// How Node::fieldScalar (Node.C:311-370; kind DNLG_H_NODE) or
// Element::getIntPointValueExtract (Element.C:1195-1221; DNLG_H_ELEMENT)
// evaluates a field: it.op (kHZero for a field the function does not know),
// it.src, it.comp.  Tensor blocks of Field.h:71-75 are norm, XX, XY, XZ, YX,
// YY, YZ, ZX, ZY, ZZ.  Returns 0, DNLG_E_INVALID (the increments StrainInc,
// PlasticStrainInc, which the engine does not keep) or DNLG_E_STATE (strain
// tensors not tracked).
int field_spec(const dnlg_job *job, int kind, int f, HistItem &it) {
  static const int tcomp[10] = {kHNorm, 0, 1, 2, 1, 3, 4, 2, 4, 5};
  it.op = kHZero;
  it.src = 0;
  it.comp = 0;
  if (kind == DNLG_H_NODE) {
    if (f == DNLG_FIELD_MASS) {
      it.op = kHNodal;
      it.src = kHMass;
      return 0;
    }
    if (f < DNLG_FIELD_DENSITY) {  // the five nodal vectors: norm, X, Y, Z
      static const int vsrc[5] = {kHCoords, kHDisp, kHDispInc, kHSpeed, kHSpeedInc};
      it.op = kHNodal;
      it.src = vsrc[f / 4];
      it.comp = f % 4 == 0 ? 3 : f % 4 - 1;
      return 0;
    }
  }
  const int ipop = kind == DNLG_H_NODE ? kHNodeIP : kHElemIP;
  static const int scalars[][2] = {{DNLG_FIELD_PLASTIC_STRAIN, kHEp},    {DNLG_FIELD_PLASTIC_STRAIN_RATE, kHDep},
                                   {DNLG_FIELD_GAMMA, kHGam},            {DNLG_FIELD_GAMMA_CUMULATE, kHGamc},
                                   {DNLG_FIELD_YIELD_STRESS, kHSy},      {DNLG_FIELD_T, kHT},
                                   {DNLG_FIELD_DENSITY, kHRho},          {DNLG_FIELD_PRESSURE, kHPres},
                                   {DNLG_FIELD_INTERNAL_ENERGY, kHEint}, {DNLG_FIELD_VON_MISES, kHVm}};
  for (auto &sc : scalars)
    if (f == sc[0]) {
      it.op = ipop;
      it.src = sc[1];
      return 0;
    }
  auto in_block = [&](int base) { return f >= base && f < base + 10; };
  if (in_block(DNLG_FIELD_STRAIN_INC) || in_block(DNLG_FIELD_PLASTIC_STRAIN_INC)) return DNLG_E_INVALID;
  int base = -1, src = 0;
  if (in_block(DNLG_FIELD_STRESS)) {
    base = DNLG_FIELD_STRESS;
    src = kHSig;
  } else if (in_block(DNLG_FIELD_STRAIN)) {
    base = DNLG_FIELD_STRAIN;
    src = kHEps;
  } else if (in_block(DNLG_FIELD_PLASTIC_STRAIN_TENSOR)) {
    base = DNLG_FIELD_PLASTIC_STRAIN_TENSOR;
    src = kHEpsP;
  }
  if (base < 0) return 0;
  if (src != kHSig && !job->track) return DNLG_E_STATE;
  it.op = ipop;
  it.src = src;
  it.comp = tcomp[f - base];
  return 0;
}

// Node -> (element, local node) lists in element order (Node::elements,
// filled as elements are created) for the node-field kernel; built once.
int ensure_node_lists(dnlg_job *job) {
  JobDev &H = job->host;
  if (H.nf_ptr) return DNLG_OK;
  dnlg_ctx *ctx = job->ctx;
  const int nn = job->nn;
  std::vector<int32_t> conn((size_t)job->NE * nn);
  HIPCHK(ctx, hipMemcpy(conn.data(), H.conn, conn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<int32_t> ptr(job->N + 1, 0);
  for (int32_t c : conn) ptr[c + 1]++;
  for (int64_t n = 0; n < job->N; n++) ptr[n + 1] += ptr[n];
  std::vector<int2> pairs(conn.size());
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t e = 0; e < job->NE; e++)
    for (int a = 0; a < nn; a++) pairs[fill[conn[e * nn + a]]++] = make_int2((int)e, a);
  int32_t *d_ptr;
  int2 *d_pairs;
  int st;
  if ((st = dalloc(job, &d_ptr, (int64_t)ptr.size(), ptr.data())) ||
      (st = dalloc(job, &d_pairs, (int64_t)pairs.size(), pairs.data())))
    return st;
  H.nf_ptr = d_ptr;
  H.nf_pairs = d_pairs;
  const size_t off = offsetof(JobDev, nf_ptr), end = offsetof(JobDev, nf_pairs) + sizeof(H.nf_pairs);
  HIPCHK(ctx, hipMemcpy((char *)job->dev + off, (const char *)&H + off, end - off, hipMemcpyHostToDevice));
  return DNLG_OK;
}

// Node::fieldScalar of every node (out [N]) with field0 in slot f0;
// tensor = true: the six components of Node::fieldSymTensor2 (out [6][N]).
int node_field(dnlg_job *job, int field, int f0, bool tensor, double *out) {
  dnlg_ctx *ctx = job->ctx;
  const int64_t N = job->N;
  const JobDev &H = job->host;
  HistItem it{};
  int st = field_spec(job, DNLG_H_NODE, field, it);
  if (st) return fail(ctx, st, "field not available at the nodes");
  if (it.op == kHZero) {
    std::memset(out, 0, (tensor ? 6 : 1) * N * sizeof(double));
    return DNLG_OK;
  }
  if (it.op == kHNodal) {
    if (it.src == kHMass) {
      HIPCHK(ctx, hipMemcpy(out, H.mass, N * sizeof(double), hipMemcpyDeviceToHost));
      return DNLG_OK;
    }
    const double *src = it.src == kHCoords || it.src == kHDisp ? H.x
                        : it.src == kHDispInc                  ? H.fu[f0]
                        : it.src == kHSpeed                    ? H.fv[f0]
                                                               : H.fa[f0];
    std::vector<double> v(3 * N, 0.0), x0;
    HIPCHK(ctx, hipMemcpy(v.data(), src, job->dim * N * sizeof(double), hipMemcpyDeviceToHost));
    if (it.src == kHDisp) {
      // the node displacement sum of Explicit.C:530-531 is x - x0
      x0.assign(3 * N, 0.0);
      HIPCHK(ctx, hipMemcpy(x0.data(), H.x0, job->dim * N * sizeof(double), hipMemcpyDeviceToHost));
      for (int64_t k = 0; k < 3 * N; k++) v[k] -= x0[k];
    }
    for (int64_t n = 0; n < N; n++)
      out[n] = it.comp == 3 ? std::sqrt(sqd(v[n]) + sqd(v[N + n]) + sqd(v[2 * N + n])) : v[it.comp * N + n];
    return DNLG_OK;
  }
  if ((st = ensure_node_lists(job))) return st;
  double *d_out = nullptr;
  const int64_t count = (tensor ? 6 : 1) * N;
  HIPCHK(ctx, hipMalloc(&d_out, count * sizeof(double)));
  launch_node_field(job->dev, N, it.src, tensor ? -1 : it.comp, d_out, ctx->stream);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d_out, count * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(d_out);
  if (e != hipSuccess) return fail(ctx, DNLG_E_HIP, std::string("node field: ") + hipGetErrorString(e));
  return DNLG_OK;
}

// VtkInterface::write (VtkInterface.C:78-278) of the job's state, field0 in
// slot f0, to <base><index, 3 digits>.vtk (DynELA::writeVTKFile,
// DynELA.C:688-710).  Numbers as C++ streams print doubles (%g).
int write_vtk(dnlg_job *job, double time, int f0) {
  dnlg_ctx *ctx = job->ctx;
  char num[32];
  std::snprintf(num, sizeof(num), "%03d", job->vtk_index);
  const std::string path = job->vtk_base + num + ".vtk";
  const int64_t N = job->N, NE = job->NE;
  const int nn = job->nn;
  std::vector<double> x(3 * N, 0.0);
  HIPCHK(ctx, hipMemcpy(x.data(), job->host.x, job->dim * N * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<int32_t> conn((size_t)NE * nn);
  HIPCHK(ctx, hipMemcpy(conn.data(), job->host.conn, conn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  FILE *fp = std::fopen(path.c_str(), "w");
  if (!fp) return fail(ctx, DNLG_E_INVALID, "cannot open result file " + path);
  // VtkInterface::headerWrite with the name and version of
  // Sources/dnlSettings/configuration.dnl
  std::fprintf(fp, "# vtk DataFile Version 4.0\n%s v.%s\nASCII\nDATASET UNSTRUCTURED_GRID\n", "DynELA FEM Code", "4.0");
  {
    char b[40];
    *put_g(b, time) = 0;
    std::fprintf(fp, "FIELD FieldData 1\nTIME 1 1 double\n%s\n", b);
  }
  // columns -> rows of `width` values, then "%g" text
  std::vector<double> aos;
  auto put_rows = [&](const double *const *cols, int width) {
    aos.resize((size_t)N * width);
    for (int64_t n = 0; n < N; n++)
      for (int c = 0; c < width; c++) aos[n * width + c] = cols[c][n];
    const std::string txt = format_rows(aos.data(), N, width, false, false);
    std::fwrite(txt.data(), 1, txt.size(), fp);
  };
  std::fprintf(fp, "POINTS %lld float\n", (long long)N);
  const double *xc[3] = {x.data(), x.data() + N, x.data() + 2 * N};
  put_rows(xc, 3);
  std::fprintf(fp, "\nCELLS %lld %lld\n", (long long)NE, (long long)(NE * (nn + 1)));
  {
    std::string txt;
    txt.reserve((size_t)NE * (nn + 1) * 8);
    char b[16];
    for (int64_t e = 0; e < NE; e++) {
      txt.append(b, std::to_chars(b, b + 16, nn).ptr - b) += ' ';
      for (int a = 0; a < nn; a++) txt.append(b, std::to_chars(b, b + 16, conn[e * nn + a]).ptr - b) += ' ';
      txt += '\n';
    }
    // vtk cell types of ElQua4NAx (9, quad) and ElHex8N3D (12, hexahedron)
    txt += "\nCELL_TYPES " + std::to_string(NE) + "\n";
    // VTK cell types (ElementData's VTK type: hexahedron 12, quad 9, tetra 10)
    const char *ct = job->type == DNLG_EL_HEX8N_3D ? "12\n" : (job->type == DNLG_EL_TET4N_3D ? "10\n" : "9\n");
    for (int64_t e = 0; e < NE; e++) txt += ct;
    std::fwrite(txt.data(), 1, txt.size(), fp);
  }
  std::fprintf(fp, "\nPOINT_DATA %lld\n", (long long)N);
  std::vector<double> v(6 * N);
  int st = DNLG_OK;
  for (int32_t f : job->vtk_fields) {
    // Field::getType (Field.C): the five nodal vectors and the five tensors
    // by their block's first value, everything else a scalar
    const bool vec = f == DNLG_FIELD_COORDS || f == DNLG_FIELD_DISP || f == DNLG_FIELD_DISP_INC ||
                     f == DNLG_FIELD_SPEED || f == DNLG_FIELD_SPEED_INC;
    const bool ten = f == DNLG_FIELD_STRESS || f == DNLG_FIELD_STRAIN || f == DNLG_FIELD_PLASTIC_STRAIN_TENSOR ||
                     f == DNLG_FIELD_STRAIN_INC || f == DNLG_FIELD_PLASTIC_STRAIN_INC;
    // fields the engine does not keep (increments, untracked strains) are left out
    HistItem probe{};
    if (field_spec(job, DNLG_H_NODE, f, probe)) continue;
    const std::string label = vtk_label(f);
    if (vec) {
      for (int d = 0; d < 3 && !st; d++) st = node_field(job, f + 1 + d, f0, false, v.data() + d * N);
      if (st) break;
      std::fprintf(fp, "VECTORS %s float\n", label.c_str());
      const double *vc[3] = {v.data(), v.data() + N, v.data() + 2 * N};
      put_rows(vc, 3);
    } else if (ten) {
      if ((st = node_field(job, f, f0, true, v.data()))) break;
      std::fprintf(fp, "TENSORS %s float\n", label.c_str());
      // t(0,0) t(0,1) t(0,2) t(1,0) t(1,1) t(1,2) t(2,0) t(2,1) t(2,2)
      static const int k9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
      const double *tc[9];
      for (int k = 0; k < 9; k++) tc[k] = v.data() + k9[k] * N;
      put_rows(tc, 9);
    } else {
      if ((st = node_field(job, f, f0, false, v.data()))) break;
      std::fprintf(fp, "SCALARS %s float\nLOOKUP_TABLE default\n", label.c_str());
      const double *sc[1] = {v.data()};
      put_rows(sc, 1);
    }
  }
  std::fprintf(fp, "\n");
  std::fclose(fp);
  if (st) return st;
  job->vtk_index++;
  return DNLG_OK;
}

}  // namespace

namespace dnlg {

// HistoryFile::save's row output (HistoryFile.C:301-312): "%10.7E " per
// value, time first, a newline per row, flushed.
int history_drain(dnlg_job *job) {
  const int nf = (int)job->hist.size();
  if (nf == 0) return DNLG_OK;
  dnlg_ctx *ctx = job->ctx;
  std::vector<HistFile> f(nf);
  HIPCHK(ctx, hipMemcpy(f.data(), job->host.hf, nf * sizeof(HistFile), hipMemcpyDeviceToHost));
  for (int i = 0; i < nf; i++) {
    dnlg_history &h = job->hist[i];
    const int64_t n_new = f[i].rows - h.drained;
    if (n_new <= 0) continue;
    if (n_new > h.cap) return fail(ctx, DNLG_E_CAPACITY, "history rows overran the device ring");
    const int64_t width = 1 + h.n_items;
    const size_t old = h.rows.size();
    h.rows.resize(old + n_new * width);
    const int64_t r0 = h.drained % h.cap;
    const int64_t first = std::min(n_new, (int64_t)h.cap - r0);
    HIPCHK(ctx, hipMemcpy(h.rows.data() + old, h.ring + r0 * width, first * width * sizeof(double),
                          hipMemcpyDeviceToHost));
    if (n_new > first)
      HIPCHK(ctx, hipMemcpy(h.rows.data() + old + first * width, h.ring, (n_new - first) * width * sizeof(double),
                            hipMemcpyDeviceToHost));
    if (h.fp) {
      const std::string txt = format_rows(h.rows.data() + old, n_new, width, true, true);
      std::fwrite(txt.data(), 1, txt.size(), h.fp);
      std::fflush(h.fp);
    }
    h.drained = f[i].rows;
  }
  return DNLG_OK;
}

// A job held at the end of a cycle: that cycle is closed (the record in
// slot 0 has its new time and the swapped field0), the file takes the
// written cycle's field0, the other slot, as the reference's endStep writes
// before its field swap.
int result_file_release(dnlg_job *job) {
  dnlg_ctx *ctx = job->ctx;
  Control c;
  HIPCHK(ctx, hipMemcpy(&c, &job->dev->c[0], sizeof(Control), hipMemcpyDeviceToHost));
  int st = write_vtk(job, c.t, 1 - c.f0);
  if (st) return st;
  const int32_t zero = 0;
  HIPCHK(ctx, hipMemcpy(&job->dev->res.hold, &zero, sizeof(zero), hipMemcpyHostToDevice));
  return DNLG_OK;
}

}  // namespace dnlg

// ===========================================================================
extern "C" {

int dnlg_abi_version(void) { return DNLG_ABI_VERSION; }

int dnlg_ctx_create(int device, dnlg_ctx **out) {
  if (!out) return DNLG_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DNLG_E_HIP;
  if (device < 0 || device >= ndev) return DNLG_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return DNLG_E_HIP;
  dnlg_ctx *ctx = new dnlg_ctx();
  ctx->device = device;
  // DNLG_GRAPH_CYCLES=<n>: cycles per captured graph; 0 launches every
  // kernel directly (for tools that do not follow graph launches)
  if (const char *g = std::getenv("DNLG_GRAPH_CYCLES")) ctx->graph_cycles = std::max(0, std::atoi(g));
  // DNLG_PI_SLOTS=<n>: power iterations per time-step cycle inside a launch
  // sequence (0..16); a job that needs more is finished by the host
  if (const char *g = std::getenv("DNLG_PI_SLOTS")) ctx->pi_slots = std::min(16, std::max(0, std::atoi(g)));
  // DNLG_BRANCHES=0: a batch's cycle kernels all on one stream
  if (const char *g = std::getenv("DNLG_BRANCHES")) ctx->branches = std::atoi(g) != 0;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return DNLG_E_HIP;
  }
  if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) != hipSuccess) {
    if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
    if (ctx->side) hipStreamDestroy(ctx->side);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return DNLG_E_HIP;
  }
  // the device address of the allocation-free entry points' scratch, once
  // per device, with the device current
  if (energy_scratch_init(device) != DNLG_OK) {
    hipEventDestroy(ctx->join_ev);
    hipEventDestroy(ctx->fork_ev);
    hipStreamDestroy(ctx->side);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return DNLG_E_HIP;
  }
  {
    std::lock_guard<std::mutex> lk(g_tables_mu);
    if (!g_tables_uploaded[device]) {
      double hexN[64], hexD[192], quaN[16], quaD[32];
      make_tables(hexN, hexD, quaN, quaD);
      upload_tables(hexN, hexD, quaN, quaD);
      upload_extrapolation();
      if (hipDeviceSynchronize() != hipSuccess) {
        hipEventDestroy(ctx->join_ev);
        hipEventDestroy(ctx->fork_ev);
        hipStreamDestroy(ctx->side);
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return DNLG_E_HIP;
      }
      g_tables_uploaded[device] = 1;
    }
  }
  *out = ctx;
  return DNLG_OK;
}

int dnlg_ctx_destroy(dnlg_ctx *ctx) {
  if (!ctx) return DNLG_E_INVALID;
  hipSetDevice(ctx->device);
  batch_cache_clear(ctx);
  hipStreamSynchronize(ctx->stream);
  if (ctx->energy_ws) hipFree(ctx->energy_ws);
  hipEventDestroy(ctx->join_ev);
  hipEventDestroy(ctx->fork_ev);
  hipStreamDestroy(ctx->side);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return DNLG_OK;
}

const char *dnlg_last_error(const dnlg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t dnlg_footprint_bytes(int elem_type, int64_t n_nodes, int64_t n_elems, int precision, int track_tensors) {
  if (!type_known(elem_type)) return -1;
  return footprint(elem_type, n_nodes, n_elems, precision, track_tensors);
}

int64_t dnlg_model_footprint_bytes(const dnlg_mesh *mesh, const dnlg_material *mat, const dnlg_bcs *bcs,
                                   const dnlg_solver_params *prm) {
  if (!mesh || !mat || !prm) return -1;
  const int type = mesh->elem_type;
  if (!type_known(type)) return -1;
  const int nn = type_nodes(type);
  int64_t n_slots = -1;
  if (mesh->conn && mesh->n_nodes > 0 && mesh->n_elems > 0) {
    bool in_range = true;
    for (int64_t k = 0; k < mesh->n_elems * nn && in_range; k++)
      in_range = mesh->conn[k] >= 0 && mesh->conn[k] < mesh->n_nodes;
    if (in_range) n_slots = incidence_slots(mesh->conn, mesh->n_nodes, mesh->n_elems, nn, nullptr, nullptr, nullptr);
  }
  int64_t fp = footprint(type, mesh->n_nodes, mesh->n_elems, prm->precision, prm->track_tensors, n_slots);
  // beyond the mesh-size estimate: the law's table (tabular point pairs or
  // the packed network) and the constant-BC entry lists
  if (mat->law == DNLG_LAW_TABULAR && mat->table) fp += 2 * mat->n_table * (int64_t)sizeof(double);
  if (mat->law == DNLG_LAW_ANN2SIG && mat->table) fp += mat->n_table * (int64_t)sizeof(double);
  if (bcs) {
    if (bcs->constant_ptr) fp += bcs->constant_ptr[mesh->n_nodes] * (int64_t)sizeof(int32_t);
    fp += (int64_t)bcs->n_boundaries * (int64_t)sizeof(Boundary);
  }
  return fp;
}

int dnlg_job_create(dnlg_ctx *ctx, const dnlg_mesh *mesh, const dnlg_material *mat, const dnlg_bcs *bcs,
                    const dnlg_solver_params *prm, dnlg_job **out) {
  if (!ctx || !mesh || !mat || !prm || !out) return fail(ctx, DNLG_E_INVALID, "null argument");
  *out = nullptr;
  if (!type_known(mesh->elem_type)) return fail(ctx, DNLG_E_INVALID, "unsupported element type");
  if (mesh->elem_type == DNLG_EL_TET4N_3D) {
    // ElTet4N3D::getCharacteristicLength returns 0 (ElTet4N3D.h:45): the
    // reference's Courant step of a tetrahedral model is 0
    if (prm->dt_method == DNLG_DT_COURANT)
      return fail(ctx, DNLG_E_INVALID,
                  "ElTet4N3D has no characteristic length (0 in the reference): use a power-iteration time step");
    if (prm->precision != DNLG_PREC_FP64)
      return fail(ctx, DNLG_E_INVALID, "ElTet4N3D runs in FP64 precision only");
  }
  if (mesh->n_nodes <= 0 || mesh->n_elems <= 0 || !mesh->coords || !mesh->conn)
    return fail(ctx, DNLG_E_INVALID, "empty mesh");
  // node, integration-point and incidence-slot indices are 32-bit (slots:
  // at most n_elems nn + n_nodes nn)
  if ((mesh->n_nodes + mesh->n_elems) * 8 >= (int64_t)1 << 31)
    return fail(ctx, DNLG_E_CAPACITY, "mesh exceeds 32-bit indexing");
  if (prm->dt_method != DNLG_DT_COURANT && !is_pi(prm->dt_method))
    return fail(ctx, DNLG_E_INVALID, "unknown time-step method");
  if (prm->stress_method != DNLG_STRESS_NR && prm->stress_method != DNLG_STRESS_DIRECT)
    return fail(ctx, DNLG_E_INVALID, "unknown stress integration method");
  if (prm->stress_method == DNLG_STRESS_DIRECT && prm->precision == DNLG_PREC_MIXED_RR32)
    return fail(ctx, DNLG_E_INVALID, "the Direct stress integration has no FP32 form (mixed-rr32)");
  if (prm->dt_freq <= 0 || !(prm->safety > 0) || prm->rho_b < 0.0 || prm->rho_b > 1.0 || prm->nr_max_iter < 0)
    return fail(ctx, DNLG_E_INVALID, "invalid solver parameters");
  if (prm->precision != DNLG_PREC_FP64 && prm->precision != DNLG_PREC_MIXED &&
      prm->precision != DNLG_PREC_MIXED_RR32)
    return fail(ctx, DNLG_E_INVALID, "invalid precision");
  if (mat->law < DNLG_LAW_JOHNSON_COOK || mat->law > DNLG_LAW_ANN2SIG)
    return fail(ctx, DNLG_E_INVALID, "invalid hardening law");
  if (mat->law == DNLG_LAW_ANN2SIG) {
    // the packed network: sizes, finite weights, non-empty scaling ranges
    const double *t = mat->table;
    if (!t || mat->n_table < 2 || !(t[0] >= 1 && t[0] <= DNLG_ANN_MAX_HIDDEN && t[0] == (int)t[0]) ||
        !(t[1] >= 1 && t[1] <= DNLG_ANN_MAX_HIDDEN && t[1] == (int)t[1]))
      return fail(ctx, DNLG_E_INVALID, "a network law needs 1 to 64 neurons per hidden layer");
    const int h1 = (int)t[0], h2 = (int)t[1];
    if (mat->n_table != DNLG_ANN_TABLE_SIZE(h1, h2))
      return fail(ctx, DNLG_E_INVALID, "network law table size does not match its layer sizes");
    for (int64_t k = 0; k < mat->n_table; k++)
      if (!std::isfinite(t[k])) return fail(ctx, DNLG_E_INVALID, "network law weights must be finite");
    const double *tail = t + mat->n_table - 9;  // logBase, min[4], max[4] (the last 9 entries)
    for (int k = 0; k < 4; k++)
      if (!(tail[5 + k] != tail[1 + k]))
        return fail(ctx, DNLG_E_INVALID, "network law scaling ranges must not be empty");
    if (!(tail[0] > 0)) return fail(ctx, DNLG_E_INVALID, "network law logBase must be > 0");
    if (prm->precision == DNLG_PREC_MIXED_RR32)
      return fail(ctx, DNLG_E_INVALID, "the network law has no FP32 radial return (mixed-rr32)");
  }
  if (mat->law == DNLG_LAW_TABULAR) {
    // DiscreteFunction with at least one segment, strains strictly ascending
    if (!mat->table || mat->n_table < 2 || mat->n_table > (1 << 20))
      return fail(ctx, DNLG_E_INVALID, "a tabular law needs 2 to 2^20 points");
    for (int64_t k = 0; k < mat->n_table; k++)
      if (!std::isfinite(mat->table[2 * k]) || !std::isfinite(mat->table[2 * k + 1]) ||
          (k > 0 && !(mat->table[2 * k] > mat->table[2 * k - 2])))
        return fail(ctx, DNLG_E_INVALID, "tabular law points must be finite with ascending strains");
    if (prm->precision == DNLG_PREC_MIXED_RR32)
      return fail(ctx, DNLG_E_INVALID, "the tabular law has no FP32 radial return (mixed-rr32)");
  }
  if (!(mat->density > 0) || !(mat->young > 0) || !(mat->heat_capacity > 0))
    return fail(ctx, DNLG_E_INVALID, "material needs density, Young modulus and heat capacity > 0");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, DNLG_E_HIP, "hipSetDevice");

  dnlg_job *job = new dnlg_job();
  job->ctx = ctx;
  const int type = mesh->elem_type;
  const int dim = type_dim(type);
  const int nn = type_nodes(type);
  const int64_t N = mesh->n_nodes, NE = mesh->n_elems, G = NE * type_ips(type);
  job->type = type;
  job->dim = dim;
  job->nn = nn;
  job->N = N;
  job->NE = NE;
  job->G = G;
  job->prec = prm->precision;
  job->track = prm->track_tensors ? 1 : 0;
  job->t_end = prm->t_end;
  auto bail = [&](int st) {
    dnlg_job_destroy(job);
    return st;
  };

  // connectivity check; the incidence slot of each (element, local node)
  std::vector<int32_t> conn(mesh->conn, mesh->conn + NE * nn);
  for (int64_t k = 0; k < NE * nn; k++)
    if (conn[k] < 0 || conn[k] >= N) return bail(fail(ctx, DNLG_E_INVALID, "connectivity index out of range"));
  std::vector<int32_t> e2slot, ovf_ptr;
  int32_t slot_w = 0;
  const int64_t n_slots = incidence_slots(conn.data(), N, NE, nn, &slot_w, &ovf_ptr, &e2slot);
  // coordinates (ElementAxi::checkLevel2, ElementAxi.C:103-126; the plane
  // element checks nothing, Element2D.C)
  std::vector<double> x(dim * N);
  for (int d = 0; d < dim; d++)
    for (int64_t n = 0; n < N; n++) x[d * N + n] = mesh->coords[d * N + n];
  if (type == DNLG_EL_QUA4N_AX)
    for (int64_t n = 0; n < N; n++)
      if (x[n] < 0.0) {
        if (x[n] > -1e-8)
          x[n] = 0.0;
        else
          return bail(fail(ctx, DNLG_E_INVALID, "negative radius in an axisymmetric model"));
      }
  // boundary conditions
  std::vector<Boundary> bnd;
  std::vector<int32_t> bc_ptr(N + 1, 0), bc_idx;
  std::vector<double> v0(dim * N, 0.0);
  if (bcs && bcs->n_boundaries > 0) {
    if (!bcs->boundaries) return bail(fail(ctx, DNLG_E_INVALID, "boundaries missing"));
    for (int i = 0; i < bcs->n_boundaries; i++) {
      const dnlg_boundary &b = bcs->boundaries[i];
      if (b.kind != DNLG_BC_RESTRAIN && b.kind != DNLG_BC_SPEED)
        return bail(fail(ctx, DNLG_E_INVALID, "unsupported boundary kind"));
      if (b.ramp < DNLG_RAMP_NONE || b.ramp > DNLG_RAMP_SOFT)
        return bail(fail(ctx, DNLG_E_INVALID, "unknown ramp function"));
      Boundary B;
      B.kind = b.kind;
      B.ramp = b.ramp;
      for (int c = 0; c < 3; c++) B.v[c] = b.value[c];
      B.rmin = b.ramp_min;
      B.rmax = b.ramp_max;
      bnd.push_back(B);
    }
    auto check_csr = [&](const int64_t *ptr, const int32_t *idx) -> bool {
      if (!ptr) return true;
      if (ptr[0] != 0) return false;
      for (int64_t n = 0; n < N; n++)
        if (ptr[n + 1] < ptr[n]) return false;
      for (int64_t k = 0; k < ptr[N]; k++)
        if (!idx || idx[k] < 0 || idx[k] >= bcs->n_boundaries) return false;
      return true;
    };
    if (!check_csr(bcs->initial_ptr, bcs->initial_idx) || !check_csr(bcs->constant_ptr, bcs->constant_idx))
      return bail(fail(ctx, DNLG_E_INVALID, "malformed boundary lists"));
    if (bcs->constant_ptr) {
      if (bcs->constant_ptr[N] >= ((int64_t)1 << 31)) return bail(fail(ctx, DNLG_E_CAPACITY, "too many BC entries"));
      for (int64_t n = 0; n <= N; n++) bc_ptr[n] = (int32_t)bcs->constant_ptr[n];
      bc_idx.assign(bcs->constant_idx, bcs->constant_idx + bcs->constant_ptr[N]);
    }
    // BoundaryCondition::applyInitial (BoundaryCondition.C:131-150): initial
    // list then constant list; BoundarySpeed::applyInitial sets the speed
    // vector and zero acceleration (BoundarySpeed.C:63-76).
    for (int64_t n = 0; n < N; n++) {
      for (int pass = 0; pass < 2; pass++) {
        const int64_t *ptr = pass == 0 ? bcs->initial_ptr : bcs->constant_ptr;
        const int32_t *idx = pass == 0 ? bcs->initial_idx : bcs->constant_idx;
        if (!ptr) continue;
        for (int64_t k = ptr[n]; k < ptr[n + 1]; k++) {
          const dnlg_boundary &b = bcs->boundaries[idx[k]];
          if (b.kind != DNLG_BC_SPEED) continue;
          for (int d = 0; d < dim; d++) v0[d * N + n] = b.value[d];
        }
      }
    }
  }
  int32_t bc_flags = 0;
  for (int32_t k : bc_idx)
    if (!bnd.empty() && bnd[k].kind == DNLG_BC_SPEED) bc_flags |= (bnd[k].ramp == DNLG_RAMP_NONE) ? 1 : 3;
  std::vector<int32_t> bcn;  // nodes with constant BCs
  for (int64_t n = 0; n < N; n++)
    if (bc_ptr[n + 1] > bc_ptr[n]) bcn.push_back((int32_t)n);
  job->n_bcn = (int32_t)bcn.size();
  if (bcn.empty()) bcn.push_back(0);
  if (bc_idx.empty()) bc_idx.push_back(0);
  if (bnd.empty()) bnd.push_back(Boundary{});

  JobDev &H = job->host;
  std::memset(&H, 0, sizeof(JobDev));
  H.type = type;
  H.dim = dim;
  H.nn = nn;
  H.nip = type_ips(type);
  H.N = N;
  H.NE = NE;
  H.G = G;
  H.prec = prm->precision;
  H.track = job->track;
  H.law = mat->law;
  H.dt_freq = prm->dt_freq;
  H.dt_method = prm->dt_method;
  H.stress_method = prm->stress_method;
  H.nr_max_it = prm->nr_max_iter > 0 ? prm->nr_max_iter : kItMax;
  H.rho0 = mat->density;
  H.young = mat->young;
  H.nu = mat->poisson;
  H.twoG = mat->young / (1.0 + mat->poisson);                 // Material.h:110-114
  H.K = mat->young / (3.0 * (1.0 - 2.0 * mat->poisson));        // Material.h:152-156
  H.heatFrac = mat->taylor_quinney / (mat->density * mat->heat_capacity);  // Element.C:438
  H.eta = mat->taylor_quinney;
  H.cp = mat->heat_capacity;
  H.T0 = mat->T0;
  for (int k = 0; k < 8; k++) H.lp[k] = mat->p[k];
  for (int k = 0; k < 8; k++) H.lpf[k] = (float)mat->p[k];
  H.inv_dTf = mat->law == DNLG_LAW_JOHNSON_COOK ? (float)(1.0 / (mat->p[6] - mat->p[7])) : 0.f;
  H.inv_rho0 = 1.0 / mat->density;
  // Explicit::computeChungHulbertIntegrationParameters (Explicit.C:86-104)
  const double rb = prm->rho_b;
  H.alphaM = (2.0 * rb - 1.0) / (1.0 + rb);
  H.beta = (5.0 - 3.0 * rb) / ((1.0 + rb) * (1.0 + rb) * (2.0 - rb));
  H.gammaCH = 1.5 - H.alphaM;
  H.omegaS = std::sqrt((12.0 * std::pow(1.0 + rb, 3) * (2.0 - rb)) /
                       (10.0 + 15.0 * rb - rb * rb + std::pow(rb, 3) - std::pow(rb, 4)));
  H.safety = prm->safety;
  for (Control &c : H.c) {
    c.t = prm->t_start;
    c.dt = 0.0;
    c.up_to = prm->t_start;
    c.inc = 0;
    c.f0 = 0;
    c.running = 1;
    c.courant_due = 0;
    c.density_due = 0;
    c.preloop = 1;
    c.pending = 0;
  }
  H.sync.st[0] = H.sync.st[1] = 0;

  int st;
#define ALLOC(ptr, count, init)                           \
  if ((st = dalloc(job, &(ptr), (count), (init))) != 0) return bail(st);
  ALLOC(H.x, dim * N, x.data());
  ALLOC(H.x0, dim * N, x.data());
  ALLOC(H.mass, N, (const double *)nullptr);
  for (int k = 0; k < 2; k++) {
    ALLOC(H.fu[k], dim * N, (const double *)nullptr);
    ALLOC(H.fv[k], dim * N, v0.data());
    ALLOC(H.fa[k], dim * N, (const double *)nullptr);
  }
  int32_t *p32;
  Boundary *pb;
  ALLOC(p32, N + 1, bc_ptr.data());
  H.bc_ptr = p32;
  ALLOC(p32, (int64_t)bc_idx.size(), bc_idx.data());
  H.bc_idx = p32;
  ALLOC(pb, (int64_t)bnd.size(), bnd.data());
  H.bnd = pb;
  ALLOC(p32, (int64_t)bcn.size(), bcn.data());
  H.bcn = p32;
  H.n_bcn = job->n_bcn;
  H.bc_flags = bc_flags;
  H.slot_w = slot_w;
  H.ovf_ptr = nullptr;
  if (!ovf_ptr.empty()) {
    ALLOC(p32, N + 1, ovf_ptr.data());
    H.ovf_ptr = p32;
  }
  ALLOC(p32, NE * nn, e2slot.data());
  H.e2slot = p32;
  ALLOC(p32, NE * nn, conn.data());
  H.conn = p32;
  ALLOC(H.elemF, n_slots * dim, (const double *)nullptr);
  job->n_cmin = (int32_t)((NE + 255) / 256);
  H.n_cmin = job->n_cmin;
  ALLOC(H.cmin, (int64_t)job->n_cmin, (const double *)nullptr);
  H.pi = PiState{0.0, 1.0, 0, 0, 0, 0, 0, 0};
  H.n_pibmax = 0;
  if (is_pi(prm->dt_method)) {
    // starting vector of Model.C:966-970 in the reference's DOF order
    // (node * dim + d), stored [dim][N]
    const int64_t ndof = (int64_t)dim * N;
    std::vector<double> ev(ndof), evs(ndof);
    ev[0] = 1.0;
    for (int64_t i = 1; i < ndof; i++) ev[i] = ev[i - 1] - 2. / (ndof - 1);
    for (int64_t n = 0; n < N; n++)
      for (int d = 0; d < dim; d++) evs[d * N + n] = ev[n * dim + d];
    ALLOC(H.pi_ev[0], ndof, evs.data());
    ALLOC(H.pi_ev[1], ndof, (const double *)nullptr);
    H.n_pibmax = (int32_t)((N + kNodeBlock - 1) / kNodeBlock);
    ALLOC(H.pi_bmax, (int64_t)H.n_pibmax, (const double *)nullptr);
  }
  std::vector<double> fillT(G, mat->T0), fillR(G, mat->density);
  ALLOC(H.ep, G, (const double *)nullptr);
  ALLOC(H.eint, G, (const double *)nullptr);
  ALLOC(H.einel, G, (const double *)nullptr);
  ALLOC(H.T, G, fillT.data());
  ALLOC(H.rho, G, fillR.data());
  ALLOC(H.detJ0, G, (const double *)nullptr);
  if (mat->law == DNLG_LAW_TABULAR) {
    std::vector<double> t(2 * mat->n_table);
    for (int64_t k = 0; k < mat->n_table; k++) {
      t[k] = mat->table[2 * k];
      t[mat->n_table + k] = mat->table[2 * k + 1];
    }
    double *pt;
    ALLOC(pt, (int64_t)t.size(), t.data());
    H.ltab = pt;
    H.n_ltab = (int32_t)mat->n_table;
  }
  if (mat->law == DNLG_LAW_ANN2SIG) {
    double *pt;
    ALLOC(pt, mat->n_table, mat->table);
    H.ltab = pt;
    H.n_ltab = (int32_t)mat->n_table;
  }
  if (is_mixed(prm->precision)) {
    ALLOC(H.elemF32, n_slots * dim, (const float *)nullptr);
    ALLOC(H.sig32, 6 * G, (const float *)nullptr);
    ALLOC(H.dep32, G, (const float *)nullptr);
    ALLOC(H.sy32, G, (const float *)nullptr);
    ALLOC(H.gam32, G, (const float *)nullptr);
    ALLOC(H.detJ0f, G, (const float *)nullptr);
  } else {
    ALLOC(H.sig, 6 * G, (const double *)nullptr);
    ALLOC(H.dep, G, (const double *)nullptr);
    ALLOC(H.sy, G, (const double *)nullptr);
    ALLOC(H.gam, G, (const double *)nullptr);
    ALLOC(H.gamc, G, (const double *)nullptr);
  }
  if (job->track) {
    ALLOC(H.eps, 6 * G, (const double *)nullptr);
    ALLOC(H.epsP, 6 * G, (const double *)nullptr);
  }
  JobDev *dev;
  ALLOC(dev, 1, &H);
  job->dev = dev;
#undef ALLOC
  *out = job;
  return DNLG_OK;
}

int dnlg_job_destroy(dnlg_job *job) {
  if (!job) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  batch_cache_forget(ctx, job);
  for (void *p : job->allocs) hipFree(p);
  for (void *p : job->hist_tables) hipFree(p);
  delete job;
  return DNLG_OK;
}

int dnlg_job_status(const dnlg_job *job) { return job ? job->status : DNLG_E_INVALID; }

int dnlg_job_solve(dnlg_job *job, double t_up_to) {
  if (!job) return DNLG_E_INVALID;
  dnlg_job *const jobs[1] = {job};
  return dnlg_batch_solve(job->ctx, jobs, 1, &t_up_to);
}

int dnlg_batch_solve(dnlg_ctx *ctx, dnlg_job *const *jobs, int n_jobs, const double *t_up_to) {
  return dnlg_batch_advance(ctx, jobs, n_jobs, t_up_to, -1);
}

int dnlg_batch_advance(dnlg_ctx *ctx, dnlg_job *const *jobs, int n_jobs, const double *t_up_to,
                       int64_t max_cycles) {
  if (!ctx || !jobs || n_jobs <= 0 || !t_up_to) return fail(ctx, DNLG_E_INVALID, "invalid batch");
  for (int i = 0; i < n_jobs; i++) {
    if (!jobs[i] || jobs[i]->ctx != ctx) return fail(ctx, DNLG_E_INVALID, "job of another context");
    for (int k = 0; k < i; k++)
      if (jobs[k] == jobs[i]) return fail(ctx, DNLG_E_INVALID, "job listed twice");
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, DNLG_E_HIP, "hipSetDevice");
  std::vector<dnlg_job *> list(jobs, jobs + n_jobs);
  std::vector<double> up(t_up_to, t_up_to + n_jobs);
  return batch_run(ctx, list, up, max_cycles);
}

int dnlg_job_sizes(const dnlg_job *job, int64_t *n_nodes, int64_t *n_elems, int32_t *n_ip, int32_t *dim) {
  if (!job) return DNLG_E_INVALID;
  if (n_nodes) *n_nodes = job->N;
  if (n_elems) *n_elems = job->NE;
  if (n_ip) *n_ip = type_ips(job->type);
  if (dim) *dim = job->dim;
  return DNLG_OK;
}

static int read_control(dnlg_job *job, Control *c) {
  HIPCHK(job->ctx, hipStreamSynchronize(job->ctx->stream));
  HIPCHK(job->ctx, hipMemcpy(c, &job->dev->c[0], sizeof(Control), hipMemcpyDeviceToHost));
  return DNLG_OK;
}

int dnlg_job_read_nodes(dnlg_job *job, double *coords, double *disp, double *speed) {
  if (!job) return DNLG_E_INVALID;
  hipSetDevice(job->ctx->device);
  Control c;
  int st = read_control(job, &c);
  if (st) return st;
  const int64_t N = job->N;
  auto get = [&](double *dst, const double *src) -> int {
    if (!dst) return DNLG_OK;
    std::memset(dst, 0, 3 * N * sizeof(double));
    HIPCHK(job->ctx, hipMemcpy(dst, src, job->dim * N * sizeof(double), hipMemcpyDeviceToHost));
    return DNLG_OK;
  };
  if ((st = get(coords, job->host.x))) return st;
  if (disp) {
    // disp = sum of the increments u = x - x0 (Explicit.C:530-534)
    std::vector<double> xs(3 * N), x0(3 * N);
    if ((st = get(xs.data(), job->host.x)) || (st = get(x0.data(), job->host.x0))) return st;
    for (int64_t k = 0; k < 3 * N; k++) disp[k] = xs[k] - x0[k];
  }
  if ((st = get(speed, job->host.fv[c.f0]))) return st;
  return DNLG_OK;
}

int dnlg_job_read_mass(dnlg_job *job, double *mass) {
  if (!job || !mass) return DNLG_E_INVALID;
  hipSetDevice(job->ctx->device);
  HIPCHK(job->ctx, hipStreamSynchronize(job->ctx->stream));
  HIPCHK(job->ctx, hipMemcpy(mass, job->host.mass, job->N * sizeof(double), hipMemcpyDeviceToHost));
  return DNLG_OK;
}

int dnlg_job_read_ip(dnlg_job *job, int field, double *out) {
  if (!job || !out) return DNLG_E_INVALID;
  hipSetDevice(job->ctx->device);
  HIPCHK(job->ctx, hipStreamSynchronize(job->ctx->stream));
  const int64_t G = job->G;
  const JobDev &H = job->host;
  const bool mixed = is_mixed(job->prec);
  auto d64 = [&](const double *src, int64_t count, double *dst) -> int {
    HIPCHK(job->ctx, hipMemcpy(dst, src, count * sizeof(double), hipMemcpyDeviceToHost));
    return DNLG_OK;
  };
  auto f32 = [&](const float *src, int64_t count, double *dst) -> int {
    std::vector<float> tmp(count);
    HIPCHK(job->ctx, hipMemcpy(tmp.data(), src, count * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < count; k++) dst[k] = tmp[k];
    return DNLG_OK;
  };
  auto stress = [&](double *dst) -> int { return mixed ? f32(H.sig32, 6 * G, dst) : d64(H.sig, 6 * G, dst); };
  switch (field) {
    case DNLG_F_PLASTIC_STRAIN: return d64(H.ep, G, out);
    case DNLG_F_PLASTIC_STRAIN_RATE: return mixed ? f32(H.dep32, G, out) : d64(H.dep, G, out);
    case DNLG_F_GAMMA: return mixed ? f32(H.gam32, G, out) : d64(H.gam, G, out);
    case DNLG_F_GAMMA_CUMULATE: {
      if (!mixed) return d64(H.gamc, G, out);
      // gammaCumulate and the plastic strain grow together from 0 (Element.C:589-598):
      // ep = sqrt(2/3) sum(gamma); the mixed path keeps ep only (FP64)
      int st = d64(H.ep, G, out);
      for (int64_t g = 0; g < G && st == DNLG_OK; g++) out[g] /= kSqrt23;
      return st;
    }
    case DNLG_F_YIELD_STRESS: return mixed ? f32(H.sy32, G, out) : d64(H.sy, G, out);
    case DNLG_F_TEMPERATURE: return d64(H.T, G, out);
    case DNLG_F_DENSITY: return d64(H.rho, G, out);
    case DNLG_F_INTERNAL_ENERGY: return d64(H.eint, G, out);
    case DNLG_F_INELASTIC_ENERGY: return d64(H.einel, G, out);
    case DNLG_F_STRESS: return stress(out);
    case DNLG_F_PRESSURE:
    case DNLG_F_VON_MISES: {
      std::vector<double> s(6 * G);
      int st = stress(s.data());
      if (st) return st;
      for (int64_t g = 0; g < G; g++) {
        const double xx = s[g], xy = s[G + g], xz = s[2 * G + g], yy = s[3 * G + g], yz = s[4 * G + g],
                     zz = s[5 * G + g];
        if (field == DNLG_F_PRESSURE)
          out[g] = (xx + yy + zz) / 3.0;
        else  // dnlVonMises (Macros.h:71-72)
          out[g] = std::sqrt((sqd(xx - yy) + sqd(xx - zz) + sqd(yy - zz) + 6 * (sqd(xy) + sqd(xz) + sqd(yz))) / 2);
      }
      return DNLG_OK;
    }
    case DNLG_F_STRAIN:
      if (!job->track) return fail(job->ctx, DNLG_E_STATE, "strain tensors not tracked (track_tensors = 0)");
      return d64(H.eps, 6 * G, out);
    case DNLG_F_PLASTIC_STRAIN_TENSOR:
      if (!job->track) return fail(job->ctx, DNLG_E_STATE, "strain tensors not tracked (track_tensors = 0)");
      return d64(H.epsP, 6 * G, out);
  }
  return fail(job->ctx, DNLG_E_INVALID, "unknown field");
}

int dnlg_job_scalars(dnlg_job *job, double *time, double *time_step, int64_t *increment) {
  if (!job) return DNLG_E_INVALID;
  hipSetDevice(job->ctx->device);
  Control c;
  int st = read_control(job, &c);
  if (st) return st;
  if (time) *time = c.t;
  if (time_step) *time_step = c.dt;
  if (increment) *increment = c.inc;
  return DNLG_OK;
}

int dnlg_job_energies(dnlg_job *job, double out[4]) {
  if (!job || !out) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  hipSetDevice(ctx->device);
  const int64_t need = 6 * (int64_t)energy_blocks(job->N, job->G) + 4;
  if (ctx->energy_ws_doubles < need) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->energy_ws) hipFree(ctx->energy_ws);
    ctx->energy_ws = nullptr;
    ctx->energy_ws_doubles = 0;
    HIPCHK(ctx, hipMalloc(&ctx->energy_ws, need * sizeof(double)));
    ctx->energy_ws_doubles = need;
  }
  double *d_out = ctx->energy_ws + (need - 4);
  launch_energies(job->dev, job->N, job->G, job->dim, type_vol_scale(job->type), ctx->energy_ws, d_out,
                  ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(out, d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DNLG_OK;
}

// ---- checkpoint / restart -------------------------------------------------------
namespace {

// Device <-> host copies of one array, FP64 on the host; `f` the FP32
// device array of a mixed job (or nullptr), `d` the FP64 one.
int ip_copy(dnlg_job *job, bool to_host, double *h, double *d, float *f, int64_t count) {
  dnlg_ctx *ctx = job->ctx;
  if (d) {
    if (to_host)
      HIPCHK(ctx, hipMemcpy(h, d, count * sizeof(double), hipMemcpyDeviceToHost));
    else
      HIPCHK(ctx, hipMemcpy(d, h, count * sizeof(double), hipMemcpyHostToDevice));
    return DNLG_OK;
  }
  std::vector<float> tmp(count);
  if (to_host) {
    HIPCHK(ctx, hipMemcpy(tmp.data(), f, count * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < count; k++) h[k] = tmp[k];
  } else {
    for (int64_t k = 0; k < count; k++) tmp[k] = (float)h[k];
    HIPCHK(ctx, hipMemcpy(f, tmp.data(), count * sizeof(float), hipMemcpyHostToDevice));
  }
  return DNLG_OK;
}

// Moves the state between the host record and the device, either way.
int state_transfer(dnlg_job *job, dnlg_state *s, bool to_host) {
  dnlg_ctx *ctx = job->ctx;
  const JobDev &H = job->host;
  const int64_t N = job->N, G = job->G;
  const int dim = job->dim;
  Control c;
  HIPCHK(ctx, hipMemcpy(&c, &job->dev->c[0], sizeof(Control), hipMemcpyDeviceToHost));
  // nodal arrays: [3][N] on the host, [dim][N] on the device
  double *nd[4] = {s->coords, s->disp_inc, s->speed, s->accel};
  double *dv[4] = {H.x, H.fu[c.f0], H.fv[c.f0], H.fa[c.f0]};
  double *dv2[4] = {nullptr, H.fu[1 - c.f0], H.fv[1 - c.f0], H.fa[1 - c.f0]};
  for (int k = 0; k < 4; k++) {
    if (to_host) {
      std::memset(nd[k], 0, 3 * N * sizeof(double));
      HIPCHK(ctx, hipMemcpy(nd[k], dv[k], dim * N * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      HIPCHK(ctx, hipMemcpy(dv[k], nd[k], dim * N * sizeof(double), hipMemcpyHostToDevice));
      if (dv2[k]) HIPCHK(ctx, hipMemcpy(dv2[k], nd[k], dim * N * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  const bool mixed = is_mixed(job->prec);
  int st;
  if ((st = ip_copy(job, to_host, s->stress, H.sig, H.sig32, 6 * G))) return st;
  if (!to_host && mixed && dim == 2) {
    // the mixed kernel of the plane and axisymmetric quads carries no xz / yz
    // stress (their kinematics produce none): it neither reads nor writes
    // those rows, which stay 0
    HIPCHK(ctx, hipMemset(H.sig32 + 2 * G, 0, G * sizeof(float)));
    HIPCHK(ctx, hipMemset(H.sig32 + 4 * G, 0, G * sizeof(float)));
  }
  double *q = s->ip_scalars;
  if ((st = ip_copy(job, to_host, q + DNLG_STATE_PLASTIC_STRAIN * G, H.ep, nullptr, G)) ||
      (st = ip_copy(job, to_host, q + DNLG_STATE_PLASTIC_STRAIN_RATE * G, H.dep, H.dep32, G)) ||
      (st = ip_copy(job, to_host, q + DNLG_STATE_GAMMA * G, H.gam, H.gam32, G)) ||
      (st = ip_copy(job, to_host, q + DNLG_STATE_YIELD_STRESS * G, H.sy, H.sy32, G)) ||
      (st = ip_copy(job, to_host, q + DNLG_STATE_TEMPERATURE * G, H.T, nullptr, G)) ||
      (st = ip_copy(job, to_host, q + DNLG_STATE_INTERNAL_ENERGY * G, H.eint, nullptr, G)) ||
      (st = ip_copy(job, to_host, q + DNLG_STATE_INELASTIC_ENERGY * G, H.einel, nullptr, G)) ||
      (st = ip_copy(job, to_host, q + DNLG_STATE_DENSITY * G, H.rho, nullptr, G)))
    return st;
  double *gc = q + DNLG_STATE_GAMMA_CUMULATE * G;
  if (!mixed) {
    if ((st = ip_copy(job, to_host, gc, H.gamc, nullptr, G))) return st;
  } else if (to_host) {
    // the mixed paths keep ep only: gammaCumulate = ep / sqrt(2/3) (Element.C:589-598)
    const double *ep = q + DNLG_STATE_PLASTIC_STRAIN * G;
    for (int64_t g = 0; g < G; g++) gc[g] = ep[g] / kSqrt23;
  }
  if (job->track && s->strain && s->plastic_strain_tensor) {
    if ((st = ip_copy(job, to_host, s->strain, H.eps, nullptr, 6 * G)) ||
        (st = ip_copy(job, to_host, s->plastic_strain_tensor, H.epsP, nullptr, 6 * G)))
      return st;
  }
  if (to_host) {
    s->time = c.t;
    s->time_step = c.dt;
    s->increment = c.inc;
  } else {
    Control n{};
    n.t = s->time;
    n.dt = s->time_step;
    n.up_to = s->time;
    n.inc = s->increment;
    n.f0 = c.f0;
    HIPCHK(ctx, hipMemcpy(&job->dev->c[0], &n, sizeof(Control), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(&job->dev->c[1], &n, sizeof(Control), hipMemcpyHostToDevice));
    job->t_now = n.t;
    job->up_to = n.t;
    job->inc_now = n.inc;
  }
  return DNLG_OK;
}

}  // namespace

int dnlg_job_read_state(dnlg_job *job, dnlg_state *state) {
  if (!job || !state || !state->coords || !state->disp_inc || !state->speed || !state->accel || !state->stress ||
      !state->ip_scalars)
    return job ? fail(job->ctx, DNLG_E_INVALID, "dnlg_job_read_state: null array") : DNLG_E_INVALID;
  hipSetDevice(job->ctx->device);
  HIPCHK(job->ctx, hipStreamSynchronize(job->ctx->stream));
  return state_transfer(job, state, true);
}

int dnlg_job_write_state(dnlg_job *job, const dnlg_state *state) {
  if (!job || !state || !state->coords || !state->disp_inc || !state->speed || !state->accel || !state->stress ||
      !state->ip_scalars)
    return job ? fail(job->ctx, DNLG_E_INVALID, "dnlg_job_write_state: null array") : DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  if (job->host.dt_method != DNLG_DT_COURANT)
    return fail(ctx, DNLG_E_INVALID, "dnlg_job_write_state: Courant time step only (the power iteration's "
                                     "iterate is not part of the state)");
  if (!(state->time_step > 0) || !(state->time >= 0) || state->increment < 0)
    return fail(ctx, DNLG_E_INVALID, "dnlg_job_write_state: invalid time, time step or increment");
  if (job->track && (!state->strain || !state->plastic_strain_tensor))
    return fail(ctx, DNLG_E_INVALID, "dnlg_job_write_state: a job with track_tensors needs the strain tensors");
  hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (dnlg_job_status(job) != DNLG_OK) return fail(ctx, DNLG_E_STATE, "dnlg_job_write_state: the job has failed");
  int st = job_setup(ctx, job);
  if (st) return fail(ctx, st, "dnlg_job_write_state: set-up of the reference Jacobian and masses failed");
  if (dnlg_job_status(job) != DNLG_OK) return dnlg_job_status(job);
  return state_transfer(job, const_cast<dnlg_state *>(state), false);
}

int dnlg_job_energies_static(dnlg_job *job, double out[4]) {
  if (!job || !out) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  const int nb = energy_blocks(job->N, job->G);
  if (nb > kEnergyScratchBlocks)
    return fail(ctx, DNLG_E_CAPACITY,
                "dnlg_job_energies_static: the job needs " + std::to_string(nb) + " reduction blocks, the static "
                "scratch holds " + std::to_string(kEnergyScratchBlocks) + " (at most " +
                std::to_string((int64_t)kEnergyScratchBlocks * kEnergyItemsPerBlock) +
                " nodes or integration points); use dnlg_job_energies");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  double *ws = energy_scratch(ctx->device);
  if (!ws) return fail(ctx, DNLG_E_STATE, "dnlg_job_energies_static: scratch of this device not set up");
  double *d_out = ws + 6 * (int64_t)kEnergyScratchBlocks;
  launch_energies(job->dev, job->N, job->G, job->dim, type_vol_scale(job->type), ws, d_out, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(out, d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return DNLG_OK;
}

// ---- result (VTK) files and node fields ---------------------------------------
int dnlg_job_set_result_files(dnlg_job *job, const char *base_name, double start_time, double end_time,
                              double increment, const int32_t *fields, int32_t n_fields) {
  if (!job) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  if (!base_name || n_fields < 0 || (n_fields > 0 && !fields)) return fail(ctx, DNLG_E_INVALID, "invalid result files");
  if (job->started) return fail(ctx, DNLG_E_STATE, "result files must be set before the first solve");
  for (int32_t k = 0; k < n_fields; k++)
    if (fields[k] < 0 || fields[k] >= DNLG_FIELD_COUNT) return fail(ctx, DNLG_E_INVALID, "unknown result field");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, DNLG_E_HIP, "hipSetDevice");
  job->vtk_base = base_name;
  job->vtk_fields.assign(fields, fields + n_fields);
  job->vtk_index = 0;
  // DynELA::setSaveTimes (DynELA.C:479-486): next starts at start
  job->host.res = ResultSched{start_time, end_time, increment, 1, 0};
  HIPCHK(ctx, hipMemcpy(&job->dev->res, &job->host.res, sizeof(ResultSched), hipMemcpyHostToDevice));
  return DNLG_OK;
}

int dnlg_job_write_result_file(dnlg_job *job, int32_t force) {
  if (!job) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  if (job->vtk_base.empty()) return fail(ctx, DNLG_E_STATE, "no result files set");
  hipSetDevice(ctx->device);
  Control c;
  int st = read_control(job, &c);
  if (st) return st;
  if (!force) {
    // DynELA::writeResultFile (DynELA.C:489-503)
    ResultSched r;
    HIPCHK(ctx, hipMemcpy(&r, &job->dev->res, sizeof(r), hipMemcpyDeviceToHost));
    if (!(c.t >= r.next && c.t <= r.end)) return DNLG_OK;
    r.next += r.inc;
    HIPCHK(ctx, hipMemcpy(&job->dev->res, &r, sizeof(r), hipMemcpyHostToDevice));
  }
  return write_vtk(job, c.t, c.f0);
}

int dnlg_job_read_node_field(dnlg_job *job, int32_t field, double *out) {
  if (!job || !out) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  if (field < 0 || field >= DNLG_FIELD_COUNT) return fail(ctx, DNLG_E_INVALID, "unknown field");
  hipSetDevice(ctx->device);
  Control c;
  int st = read_control(job, &c);
  if (st) return st;
  return node_field(job, field, c.f0, false, out);
}

// ---- history files (HistoryFile.C) ------------------------------------------
int dnlg_job_add_history(dnlg_job *job, const char *path, double start_time, double save_time,
                         const dnlg_history_item *items, int32_t n_items, const char *const *names,
                         int32_t *file_id) {
  if (!job) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  if (job->started) return fail(ctx, DNLG_E_STATE, "history files must be added before the first solve");
  if (n_items < 0 || (n_items > 0 && !items)) return fail(ctx, DNLG_E_INVALID, "invalid history items");
  if (!(save_time >= 0.0) || !std::isfinite(start_time))
    return fail(ctx, DNLG_E_INVALID, "invalid history save times");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, DNLG_E_HIP, "hipSetDevice");
  const int fid = (int)job->hist.size();
  const int nip = type_ips(job->type);
  std::vector<HistItem> add;
  std::vector<int64_t> ip_nodes;  // nodes of node-from-IP items, per item (-1: none)
  bool need_ke = false, need_rho = false;
  for (int32_t k = 0; k < n_items; k++) {
    const dnlg_history_item &in = items[k];
    HistItem it{};
    it.file = fid;
    it.col = 1 + k;
    it.op = kHZero;
    it.ip = -1;
    it.index = in.index;
    const int f = in.field;
    if (f < 0 || f >= DNLG_FIELD_COUNT) return fail(ctx, DNLG_E_INVALID, "unknown history field");
    int64_t ipn = -1;
    if (in.kind == DNLG_H_NODE) {
      if (in.index < 0 || in.index >= job->N) return fail(ctx, DNLG_E_INVALID, "history node out of range");
    } else if (in.kind == DNLG_H_ELEMENT) {
      if (in.index < 0 || in.index >= job->NE) return fail(ctx, DNLG_E_INVALID, "history element out of range");
      if (in.int_pt > nip) return fail(ctx, DNLG_E_INVALID, "history integration point out of range");
      // HistoryFile.C:247: point intPt - 1; Element.C:1229: a negative one is the mean
      it.ip = in.int_pt > 0 ? in.int_pt - 1 : -1;
    }
    if (in.kind == DNLG_H_NODE || in.kind == DNLG_H_ELEMENT) {
      const int r = field_spec(job, in.kind, f, it);
      if (r) return fail(ctx, r, r == DNLG_E_STATE ? "strain tensors not tracked (track_tensors = 0)"
                                                   : "StrainInc and PlasticStrainInc are not kept by the engine");
      if (it.op == kHNodeIP) ipn = in.index;
    } else if (in.kind == DNLG_H_GLOBAL) {
      if (f == DNLG_FIELD_TIME_STEP || f == DNLG_FIELD_REAL_TIME_STEP) it.op = kHDt;
      if (f == DNLG_FIELD_KINETIC_ENERGY) it.op = kHKe;
    } else {
      return fail(ctx, DNLG_E_INVALID, "unknown history item kind");
    }
    need_ke = need_ke || it.op == kHKe;
    need_rho = need_rho || ((it.op == kHNodeIP || it.op == kHElemIP) && it.src == kHRho);
    add.push_back(it);
    ip_nodes.push_back(ipn);
  }
  // incidence lists of the nodes of node-from-IP items, in element order
  std::vector<int2> inc = job->hist_inc;
  bool any_ipn = false;
  for (int64_t n : ip_nodes) any_ipn = any_ipn || n >= 0;
  if (any_ipn) {
    const int nn = job->nn;
    std::vector<int32_t> conn((size_t)job->NE * nn);
    HIPCHK(ctx, hipMemcpy(conn.data(), job->host.conn, conn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::map<int64_t, std::vector<int2>> lists;
    for (int64_t n : ip_nodes)
      if (n >= 0) lists[n];
    for (int64_t e = 0; e < job->NE; e++)
      for (int a = 0; a < nn; a++) {
        auto itl = lists.find(conn[e * nn + a]);
        if (itl != lists.end()) itl->second.push_back(make_int2((int)e, a));
      }
    for (size_t k = 0; k < add.size(); k++) {
      if (ip_nodes[k] < 0) continue;
      const auto &l = lists[ip_nodes[k]];
      if (inc.size() + l.size() >= ((size_t)1 << 31)) return fail(ctx, DNLG_E_CAPACITY, "history lists too long");
      add[k].r0 = (int32_t)inc.size();
      inc.insert(inc.end(), l.begin(), l.end());
      add[k].r1 = (int32_t)inc.size();
    }
  }
  if (job->hist_items.size() + add.size() >= ((size_t)1 << 31)) return fail(ctx, DNLG_E_CAPACITY, "too many items");
  // ring: room for the rows of the cycles between two host drains (a few
  // captured sequences), 256 rows unless they are large
  dnlg_history h;
  h.n_items = n_items;
  const int64_t row_bytes = 8 * (1 + (int64_t)n_items);
  const int64_t base = 4 * (int64_t)std::max({ctx->graph_cycles, job->host.dt_freq, 1}) + 4;
  h.cap = (int32_t)std::max(base, std::min<int64_t>(256, std::max<int64_t>(1, ((int64_t)64 << 20) / row_bytes)));
  int st;
  if ((st = dalloc(job, &h.ring, (int64_t)h.cap * (1 + n_items), (const double *)nullptr))) return st;
  if (need_ke && !job->host.hke)
    if ((st = dalloc(job, &job->host.hke, (int64_t)hist_blocks(job->N), (const double *)nullptr))) return st;
  HistFile hf{};
  // HistoryFile::setSaveTime (HistoryFile.C:131-148): next starts at start
  hf.start = start_time;
  hf.next = start_time;
  hf.save = save_time;
  hf.rows = 0;
  hf.ring = h.ring;
  hf.cap = h.cap;
  hf.n_items = n_items;
  hf.ke0 = (int32_t)job->hist_ke_cols.size();
  std::vector<int32_t> ke_cols = job->hist_ke_cols;
  for (const HistItem &it : add)
    if (it.op == kHKe) ke_cols.push_back(it.col);
  hf.n_ke = (int32_t)ke_cols.size() - hf.ke0;
  if (path) {
    // HistoryFile::headerWrite (HistoryFile.C:340-358) with the program name
    // and version of Sources/dnlSettings/configuration.dnl
    h.path = path;
    h.fp = std::fopen(path, "w");
    if (!h.fp) return fail(ctx, DNLG_E_INVALID, std::string("cannot open history file ") + path);
    std::fprintf(h.fp, "#%s v. %s history file\n", "DynELA FEM Code", "4.0");
    std::fprintf(h.fp, "#plotted :");
    for (int32_t k = 0; k < n_items; k++) std::fprintf(h.fp, "%s ", names && names[k] ? names[k] : "");
    std::fprintf(h.fp, "\n");
    std::fflush(h.fp);
  }
  // new device tables: files, items, incidence pairs, kinetic-energy columns
  std::vector<HistFile> files = job->hist_files;
  files.push_back(hf);
  std::vector<HistItem> all = job->hist_items;
  all.insert(all.end(), add.begin(), add.end());
  std::vector<void *> tables;
  auto upload = [&](const void *src, size_t bytes, void **dst) -> int {
    *dst = nullptr;
    if (bytes == 0) return DNLG_OK;
    HIPCHK(ctx, hipMalloc(dst, bytes));
    tables.push_back(*dst);
    HIPCHK(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return DNLG_OK;
  };
  void *d_files, *d_items, *d_inc, *d_ke;
  if ((st = upload(files.data(), files.size() * sizeof(HistFile), &d_files)) ||
      (st = upload(all.data(), all.size() * sizeof(HistItem), &d_items)) ||
      (st = upload(inc.data(), inc.size() * sizeof(int2), &d_inc)) ||
      (st = upload(ke_cols.data(), ke_cols.size() * sizeof(int32_t), &d_ke))) {
    for (void *p : tables) hipFree(p);
    if (h.fp) std::fclose(h.fp);
    return st;
  }
  JobDev &H = job->host;
  H.hf = (HistFile *)d_files;
  H.hi = (const HistItem *)d_items;
  H.hinc = (const int2 *)d_inc;
  H.hke_col = (const int32_t *)d_ke;
  H.n_hf = (int32_t)files.size();
  H.n_hi = (int32_t)all.size();
  H.density_always = H.density_always || need_rho;
  // the history tail of the device record (no kernel has run on the job)
  const size_t off = offsetof(JobDev, hf);
  HIPCHK(ctx, hipMemcpy((char *)job->dev + off, (const char *)&H + off, sizeof(JobDev) - off, hipMemcpyHostToDevice));
  for (void *p : job->hist_tables) hipFree(p);
  job->hist_tables = tables;
  job->hist_files = files;
  job->hist_items = all;
  job->hist_inc = inc;
  job->hist_ke_cols = ke_cols;
  job->hist.push_back(std::move(h));
  if (file_id) *file_id = fid;
  return DNLG_OK;
}

int dnlg_job_history_rows(dnlg_job *job, int32_t file_id, double *rows, int64_t max_rows, int64_t *n_rows) {
  if (!job || file_id < 0 || file_id >= (int32_t)job->hist.size()) return DNLG_E_INVALID;
  const dnlg_history &h = job->hist[file_id];
  const int64_t width = 1 + h.n_items;
  const int64_t n = (int64_t)h.rows.size() / width;
  if (n_rows) *n_rows = n;
  if (rows && max_rows > 0) std::memcpy(rows, h.rows.data(), std::min(n, max_rows) * width * sizeof(double));
  return DNLG_OK;
}

int dnlg_ctx_set_timing(dnlg_ctx *ctx, int enable) {
  if (!ctx) return DNLG_E_INVALID;
  ctx->timing = enable != 0;
  return DNLG_OK;
}

int dnlg_ctx_kernel_time(dnlg_ctx *ctx, const char *name, double *total_ms, int64_t *launches) {
  if (!ctx || !name) return DNLG_E_INVALID;
  auto it = ctx->ktime.find(name);
  if (total_ms) *total_ms = it == ctx->ktime.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == ctx->ktime.end() ? 0 : it->second.second;
  return DNLG_OK;
}

int dnlg_ctx_reset_timing(dnlg_ctx *ctx) {
  if (!ctx) return DNLG_E_INVALID;
  ctx->ktime.clear();
  return DNLG_OK;
}

}  // extern "C"
