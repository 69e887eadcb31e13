/*
 * dnlg.h -- C ABI of the MI355X explicit-dynamics engine (libdnlg.so).
 *
 * This is the drop-in boundary for DynELA's explicit time-step solve.  In
 * the reference the solve is reached through the virtual
 *   Solver::solve(double solveUpToTime)          Sources/dnlFEM/Solver.h:70
 * called by Model::solve (Sources/dnlFEM/Model.C:696-730), which dispatches
 * to Explicit::solve (Sources/dnlFEM/Explicit.C:126-253).  Here the model is
 * plain data (mesh, material, boundary conditions, solver parameters) handed
 * over once at job creation; the engine keeps the state resident in HBM and
 * advances it on the GPU.
 *
 * Conventions
 *  - every entry point returns an int status (0 = DNLG_OK, < 0 = error);
 *    nothing ever calls exit(), unlike the reference's fatalError
 *    (Sources/dnlKernel/Errors.C:157,196);
 *  - host arrays passed in are copied at create time; read-back buffers are
 *    caller-allocated;
 *  - nodal arrays are structure-of-arrays: component d of node i sits at
 *    [d * n_nodes + i];
 *  - integration-point arrays are element-major: IP k of element e sits at
 *    [e * n_ip + k] (tensor components add an outer stride of n_elems*n_ip);
 *  - symmetric tensors use the reference's component order
 *    [xx, xy, xz, yy, yz, zz] (Sources/dnlMaths/SymTensor2.C:139-148).
 */
#ifndef DNLG_H
#define DNLG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNLG_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
#define DNLG_OK 0
#define DNLG_E_INVALID (-1)      /* bad argument or unsupported model       */
#define DNLG_E_OOM (-2)          /* device memory exhausted                 */
#define DNLG_E_NEG_JACOBIAN (-3) /* detJ < 0 (Element3D.C:98-102)           */
#define DNLG_E_NR_DIVERGED (-4)  /* radial return > 250 its (Element.C:560) */
#define DNLG_E_NONFINITE (-5)    /* NaN/Inf in the nodal state              */
#define DNLG_E_HIP (-6)          /* HIP runtime error                       */
#define DNLG_E_STATE (-7)        /* call not valid in the job's state       */
#define DNLG_E_CAPACITY (-8)     /* request exceeds a fixed capacity        */
#define DNLG_E_PI_DIVERGED (-9)  /* power iteration > 1000 its (Model.C:1002) */
#define DNLG_E_TABLE_RANGE (-10) /* plastic strain outside a tabular law's table
                                    (DiscreteFunction.C:195-196, 291-292) */

/* ---- enumerations -------------------------------------------------------- */
/* element types: values of the reference's element enum (Element.h:62-71) */
#define DNLG_EL_QUA4N_2D 1 /* 4-node plane-strain quad, 2x2 Gauss points (ElQua4N2D.C) */
#define DNLG_EL_QUA4N_AX 3 /* 4-node axisymmetric quad, 2x2 Gauss points */
#define DNLG_EL_HEX8N_3D 4 /* 8-node hexahedron, 2x2x2 Gauss points     */
#define DNLG_EL_TET4N_3D 5 /* 4-node tetrahedron, one integration point (ElTet4N3D.C);
                              FP64 precision, power-iteration time steps only: the
                              reference's characteristic length of this element is 0
                              (ElTet4N3D.h:45), so its Courant step is 0 */

/* hardening laws */
#define DNLG_LAW_JOHNSON_COOK 0 /* JohnsonCookLaw.C:146-170,298-339; p = A,B,C,n,m,depsp0,Tm,T0 */
#define DNLG_LAW_POWER 1        /* ElastoplasticLaw.C:110-121;          p = A,B,n            */
#define DNLG_LAW_VOCE_LINEAR 2  /* sy = A + Q(1-exp(-b ep)) + H ep;     p = A,Q,b,H          */
#define DNLG_LAW_TABULAR 3      /* TabularLaw.C: sy(ep) piecewise linear through the
                                   material's table (DiscreteFunction::getValue /
                                   getSlope, DiscreteFunction.C:188-308); FP64 and
                                   mixed (an FP64 radial return), not mixed-rr32 */
#define DNLG_LAW_ANN2SIG 4      /* ANN2SigLaw.C:173-245: sy(ep, dep, T) from a network of
                                   3 inputs, two sigmoid hidden layers of h1 and h2
                                   neurons and one linear output; FP64 and mixed
                                   (an FP64 radial return), not mixed-rr32.  The
                                   material's table holds the network, packed as
                                   h1, h2, w1[h1][3], b1[h1], w2[h2][h1], b2[h2],
                                   w3[h2], b3, logBase, minEntries[4], maxEntries[4]
                                   (the reference's matrices after its transposes,
                                   ANN2SigLaw.C:137-158, row-major) */
#define DNLG_ANN_MAX_HIDDEN 64  /* neurons per hidden layer, at most */
#define DNLG_ANN_TABLE_SIZE(h1, h2) (12 + 4 * (h1) + (h1) * (h2) + 2 * (h2))

/* boundary kinds (Boundary.h:59-62) */
#define DNLG_BC_RESTRAIN 0 /* BoundaryRestrain.C:98-150 */
#define DNLG_BC_SPEED 1    /* BoundarySpeed.C:63-76,139-226 */

/* ramp functions: values of RampFunction's enum (RampFunction.h:31-39) */
#define DNLG_RAMP_NONE 0
#define DNLG_RAMP_LINEAR 1
#define DNLG_RAMP_SINUS 2
#define DNLG_RAMP_CONSTANT 3
#define DNLG_RAMP_ERF 4
#define DNLG_RAMP_SOFT 5

/* time-step methods (Solver.h:28-33, Solver.C:196-236): Courant, and the
 * power iteration on the fully integrated or (PowerIterationUnder) the
 * one-point under-integrated element stiffness (Model.C:942-1009) */
#define DNLG_DT_COURANT 0
#define DNLG_DT_POWER_ITERATION 1
#define DNLG_DT_POWER_ITERATION_UNDER 2

/* stress integration (Model.h StressIntNR / StressIntDirect, Model.C:840-851) */
#define DNLG_STRESS_NR 0     /* Newton-Raphson radial return, Element.C:407-622 */
#define DNLG_STRESS_DIRECT 1 /* one linearised return, Element.C:625-768 (FP64 and
                                mixed, in FP64; not mixed-rr32) */

/* arithmetic of the element phases */
#define DNLG_PREC_FP64 0  /* everything in FP64                              */
#define DNLG_PREC_MIXED 1 /* FP32 geometry, strain, stress and internal
                             forces; FP64 radial-return Newton solve for the
                             plastic multiplier (the reference's control flow
                             and 1e-8 tolerance, Element.C:485-601), nodal
                             kinematics, accumulated plastic strain,
                             temperature and energies                         */
#define DNLG_PREC_MIXED_RR32 2 /* DNLG_PREC_MIXED with the radial return in
                             FP32 as well (Newton to 4 ulp of gamma, hardware
                             log2 / exp2 in the law): faster, looser plastic
                             strain; no counterpart in the reference          */

/* integration-point fields for dnlg_job_read_ip (names of Field.h:46-79) */
#define DNLG_F_PLASTIC_STRAIN 0      /* scalar */
#define DNLG_F_PLASTIC_STRAIN_RATE 1 /* scalar */
#define DNLG_F_GAMMA 2               /* scalar */
#define DNLG_F_GAMMA_CUMULATE 3      /* scalar */
#define DNLG_F_YIELD_STRESS 4        /* scalar */
#define DNLG_F_TEMPERATURE 5         /* scalar */
#define DNLG_F_DENSITY 6             /* scalar */
#define DNLG_F_PRESSURE 7            /* scalar, tr(Stress)/3 */
#define DNLG_F_INTERNAL_ENERGY 8     /* scalar, specific */
#define DNLG_F_INELASTIC_ENERGY 9    /* scalar, specific */
#define DNLG_F_VON_MISES 10          /* scalar */
#define DNLG_F_STRESS 11             /* 6 components */
#define DNLG_F_STRAIN 12             /* 6 components (needs track_tensors) */
#define DNLG_F_PLASTIC_STRAIN_TENSOR 13 /* 6 components (needs track_tensors) */
#define DNLG_F_COUNT 14

/* ---- model descriptors --------------------------------------------------- */
typedef struct dnlg_mesh {
  int32_t elem_type; /* DNLG_EL_* */
  int32_t reserved;
  int64_t n_nodes;
  int64_t n_elems;
  const double *coords; /* [3 * n_nodes] SoA; z ignored for axisymmetric */
  const int32_t *conn;  /* [n_elems * nodes_per_elem], 0-based node indices,
                           node order as ElQua4NAx.C:30-60 / ElHex8N3D.C:35-90 */
} dnlg_mesh;

typedef struct dnlg_material {
  double young;          /* Material::youngModulus   (Material.h:49) */
  double poisson;        /* Material::poissonRatio   (Material.h:50) */
  double density;        /* Material::density        (Material.h:51) */
  double heat_capacity;  /* Material::heatCapacity   (Material.h:52) */
  double taylor_quinney; /* Material::taylorQuinney  (Material.h:56) */
  double T0;             /* Material::T0             (Material.h:54) */
  int32_t law;           /* DNLG_LAW_* */
  int32_t reserved;
  double p[8];           /* law parameters, see DNLG_LAW_* */
  const double *table;   /* DNLG_LAW_TABULAR: n_table (plastic strain, yield
                            stress) pairs, strains ascending; DNLG_LAW_ANN2SIG:
                            the packed network, n_table doubles; else NULL */
  int64_t n_table;
} dnlg_material;

typedef struct dnlg_boundary {
  int32_t kind;     /* DNLG_BC_* */
  int32_t ramp;     /* DNLG_RAMP_* (speed only; NONE = constant speed) */
  double value[3];  /* restrain flags (> 0 = restrained) or speed vector */
  double ramp_min;  /* RampFunction::setFunction(type, min, max) */
  double ramp_max;
} dnlg_boundary;

/* Per-node boundary lists in CSR form, applied in list order exactly as the
 * reference applies a node's initial then constant lists
 * (BoundaryCondition.C:131-190).  Either CSR may be NULL (no entries). */
typedef struct dnlg_bcs {
  int32_t n_boundaries;
  int32_t reserved;
  const dnlg_boundary *boundaries;
  const int64_t *initial_ptr;  /* [n_nodes + 1] */
  const int32_t *initial_idx;  /* indices into boundaries */
  const int64_t *constant_ptr; /* [n_nodes + 1] */
  const int32_t *constant_idx;
} dnlg_bcs;

typedef struct dnlg_solver_params {
  double t_start;        /* Solver::setTimes (Solver.C) */
  double t_end;
  double rho_b;          /* Explicit::setDissipation, default 0.8182 (Explicit.h:16) */
  double safety;         /* Solver::_timeStepSafetyFactor, default 0.9 (Solver.h:30) */
  int32_t dt_method;     /* DNLG_DT_*, default Courant */
  int32_t dt_freq;       /* Solver::_computeTimeStepFrequency, default 10 */
  int32_t stress_method; /* DNLG_STRESS_NR */
  int32_t precision;     /* DNLG_PREC_* */
  int32_t track_tensors; /* 1: carry Strain and PlasticStrain tensors */
  int32_t nr_max_iter;   /* iterations of the radial return's Newton solve
                            before DNLG_E_NR_DIVERGED; 0 = the reference's
                            250 (Element.C:560-581) */
} dnlg_solver_params;

/* ---- opaque handles ------------------------------------------------------- */
typedef struct dnlg_ctx dnlg_ctx;
typedef struct dnlg_job dnlg_job;

/* ---- context --------------------------------------------------------------- */
/* One context per device.  Several contexts (and many jobs per context) may
 * live in one process: the reference allowed a single global DynELA object
 * (DynELA.C:21,47-52); this ABI keeps no global mutable state. */
int dnlg_abi_version(void);
int dnlg_ctx_create(int device, dnlg_ctx **out);
int dnlg_ctx_destroy(dnlg_ctx *ctx);
const char *dnlg_last_error(const dnlg_ctx *ctx);

/* ---- jobs (one job = one mesh) ---------------------------------------------- */
/* Model::initSolve (Model.C:360-521): topology and axisymmetric radius
 * checks, IP initialisation (T = T0, rho = rho0), initial boundary
 * conditions.  Copies every host array. */
int dnlg_job_create(dnlg_ctx *ctx, const dnlg_mesh *mesh,
                    const dnlg_material *mat, const dnlg_bcs *bcs,
                    const dnlg_solver_params *params, dnlg_job **out);
int dnlg_job_destroy(dnlg_job *job);

/* Replaces Solver::solve(double) (Solver.h:70) -> Explicit::solve
 * (Explicit.C:126-253): advances the job until its time reaches t_up_to
 * (clipped to t_end).  The pre-loop block (reference Jacobian, lumped mass,
 * forced time step, initial strain/stress/force pass) runs on the first
 * call only.  Blocking. */
int dnlg_job_solve(dnlg_job *job, double t_up_to);

/* Advances several jobs of one context together (one launch per phase and
 * cycle covers all of them); t_up_to[i] per job.  Blocking. */
int dnlg_batch_solve(dnlg_ctx *ctx, dnlg_job *const *jobs, int n_jobs,
                     const double *t_up_to);

/* dnlg_batch_solve that returns after about max_cycles cycles (whole
 * graphs) even if some job has not reached its target; read each job's time
 * with dnlg_job_scalars.  A later call continues where it stopped.  No
 * counterpart in the reference, whose loop runs to the target time. */
int dnlg_batch_advance(dnlg_ctx *ctx, dnlg_job *const *jobs, int n_jobs,
                       const double *t_up_to, int64_t max_cycles);

/* ---- serving loop ----------------------------------------------------------
 * The reference solves one model per process to its end time (Model::solve,
 * Model.C:696-730, through Solver::solve, Solver.h:70); it has no queue.
 * The server is that same solve run as a service: submitted models wait in a
 * queue, are created on the device when their footprint
 * (dnlg_footprint_bytes) fits in the memory budget, and every admitted job
 * advances toward its t_end together with the others, max_cycles cycles per
 * dnlg_server_step (continuous batching: jobs join and leave between steps).
 * A job's results do not depend on the other jobs it ran with. */
typedef struct dnlg_server dnlg_server;
/* hbm_budget <= 0: 85 % of the device memory free at creation;
 * max_jobs <= 0: 1024 jobs in flight at most */
int dnlg_server_create(dnlg_ctx *ctx, int64_t hbm_budget, int max_jobs,
                       dnlg_server **out);
/* destroys every job the server still holds */
int dnlg_server_destroy(dnlg_server *srv);
/* copies the descriptors (as dnlg_job_create would) and queues the job;
 * DNLG_E_CAPACITY if its footprint exceeds the whole budget */
int dnlg_server_submit(dnlg_server *srv, const dnlg_mesh *mesh,
                       const dnlg_material *mat, const dnlg_bcs *bcs,
                       const dnlg_solver_params *params, int64_t *ticket);
/* admits what fits, advances the running jobs by about max_cycles cycles,
 * retires finished and failed ones.  Blocking. */
int dnlg_server_step(dnlg_server *srv, int max_cycles, int *n_running,
                     int *n_queued);
/* cycles > 0: a job still short of t_end after that many increments is
 * stopped and finishes with DNLG_E_CAPACITY (its state can still be read
 * after dnlg_server_take); 0: no limit (default) */
int dnlg_server_set_cycle_budget(dnlg_server *srv, int64_t cycles);
/* state: 0 queued, 1 running, 2 finished; status: the job's status code */
int dnlg_server_poll(dnlg_server *srv, int64_t ticket, int *state, int *status);
/* hands a finished job over to the caller (read it with dnlg_job_read_*, free
 * it with dnlg_job_destroy) and forgets the ticket; returns the job's final
 * status.  *job is NULL when the job could not be created. */
int dnlg_server_take(dnlg_server *srv, int64_t ticket, dnlg_job **job);

/* ---- read-back ------------------------------------------------------------ */
int dnlg_job_sizes(const dnlg_job *job, int64_t *n_nodes, int64_t *n_elems,
                   int32_t *n_ip, int32_t *dim);
/* any pointer may be NULL; each is [3 * n_nodes] */
int dnlg_job_read_nodes(dnlg_job *job, double *coords, double *disp,
                        double *speed);
/* out: [ncomp * n_elems * n_ip], ncomp = 6 for tensor fields, else 1 */
int dnlg_job_read_ip(dnlg_job *job, int field, double *out);
/* time, current time step, increment count, cycles run by the engine */
int dnlg_job_scalars(dnlg_job *job, double *time, double *time_step,
                     int64_t *increment);
/* out[0] kinetic energy (Model.C:612-625), out[1] internal (strain)
 * energy sum_ip rho0 w detJ0 e_int, out[2] dissipated plastic work
 * sum_ip rho0 w detJ0 e_inel, out[3] out[0] + out[1] */
int dnlg_job_energies(dnlg_job *job, double out[4]);
/* dnlg_job_energies without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (tight serving
 * loops, code that forbids hipMalloc).  Same kernels, same bits as
 * dnlg_job_energies; the reduction's partials go to a fixed scratch array
 * of 192 KiB per GPU held until the process ends, so the job may have at
 * most 4096 x 4096 = 16 777 216 nodes and as many integration points
 * (DNLG_E_CAPACITY beyond).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (several threads or contexts on the same
 * device) race and must be serialised by the caller. */
int dnlg_job_energies_static(dnlg_job *job, double out[4]);
/* lumped nodal masses, [n_nodes] (Model.C:537-590) */
int dnlg_job_read_mass(dnlg_job *job, double *mass);
/* device bytes a job of this size and options holds (admission control) */
int64_t dnlg_footprint_bytes(int elem_type, int64_t n_nodes, int64_t n_elems,
                             int precision, int track_tensors);
/* dnlg_footprint_bytes plus what the descriptors add: the law's table and
 * the constant-BC entry lists; the figure the server admits a job by
 * (-1 for an unsupported element type) */
int64_t dnlg_model_footprint_bytes(const dnlg_mesh *mesh, const dnlg_material *mat, const dnlg_bcs *bcs,
                                   const dnlg_solver_params *params);
int dnlg_job_status(const dnlg_job *job);

/* ---- checkpoint / restart --------------------------------------------------
 * The solver state of a job as plain host arrays.  The reference keeps no
 * checkpoint (Model::readData / writeData are commented out,
 * Model.C:1031-1100); its dormant NodalField::write / read
 * (NodalField.C:97-130) and IntegrationPoint::write / read
 * (IntegrationPoint.C:142-171) cover the same fields.  A job whose state is
 * written continues exactly as the job it was read from would (FP64: the
 * same bits), so a solve can be stopped, saved and resumed elsewhere -- on
 * another device, or in the CPU oracle.  All arrays are caller-allocated;
 * nodal arrays [3 * n_nodes] SoA (2D: z components ignored / zero),
 * integration-point arrays element-major [n_elems * n_ip] per component. */
#define DNLG_STATE_PLASTIC_STRAIN 0      /* rows of dnlg_state.ip_scalars */
#define DNLG_STATE_PLASTIC_STRAIN_RATE 1
#define DNLG_STATE_GAMMA 2
#define DNLG_STATE_GAMMA_CUMULATE 3      /* mixed paths: derived from the plastic strain, ignored on write */
#define DNLG_STATE_YIELD_STRESS 4
#define DNLG_STATE_TEMPERATURE 5
#define DNLG_STATE_INTERNAL_ENERGY 6
#define DNLG_STATE_INELASTIC_ENERGY 7
#define DNLG_STATE_DENSITY 8
#define DNLG_STATE_IP_SCALARS 9
typedef struct dnlg_state {
  double time;          /* model->currentTime */
  double time_step;     /* Solver::timeStep (the step the next cycle starts with) */
  int64_t increment;    /* Solver::currentIncrement */
  double *coords;       /* current coordinates */
  double *disp_inc;     /* field0 u: the last cycle's displacement increment (NodalField.h:32) */
  double *speed;        /* field0 v */
  double *accel;        /* field0 a */
  double *stress;       /* [6][n_elems * n_ip] */
  double *ip_scalars;   /* [DNLG_STATE_IP_SCALARS][n_elems * n_ip] */
  double *strain;       /* [6][n_elems * n_ip] Strain, track_tensors only (else NULL) */
  double *plastic_strain_tensor; /* [6][n_elems * n_ip], track_tensors only (else NULL) */
} dnlg_state;
/* Copies the state of a job that is between solves. */
int dnlg_job_read_state(dnlg_job *job, dnlg_state *state);
/* Sets the state of a job (Courant time step only; not after a failure).
 * A job that has not been solved yet first takes its reference Jacobian and
 * lumped masses from its mesh, as the pre-loop does (Model.C:754-775,
 * 537-590); the pre-loop's forced time step and element pass are skipped:
 * the state stands for them.  The job then continues from state->time. */
int dnlg_job_write_state(dnlg_job *job, const dnlg_state *state);

/* ---- history files --------------------------------------------------------
 * HistoryFile (Sources/dnlFEM/HistoryFile.C) attached to a model with
 * Model::add(HistoryFile *) (Model.C:927-939).  A file takes a row -- the
 * time, then every item in the order added, each printed "%10.7E " -- after
 * the pre-loop pass and at the end of every cycle whose new time t has
 * t <= start or t >= next; a row moves next one save interval on
 * (HistoryFile::save, HistoryFile.C:283-315).  setSaveTime(save) is
 * start = 0, setSaveTime(start, stop, save) is start (the reference does not
 * use the stop time either).  Item fields are the values of the reference's
 * Field::FieldLabel enum (Sources/dnlKernel/Field.h:46-79); a vector block is
 * norm, X, Y, Z and a tensor block norm, XX, XY, XZ, YX, YY, YZ, ZX, ZY, ZZ. */
#define DNLG_FIELD_COORDS 0
#define DNLG_FIELD_DISP 4
#define DNLG_FIELD_DISP_INC 8
#define DNLG_FIELD_SPEED 12
#define DNLG_FIELD_SPEED_INC 16
#define DNLG_FIELD_DENSITY 20
#define DNLG_FIELD_GAMMA 21
#define DNLG_FIELD_GAMMA_CUMULATE 22
#define DNLG_FIELD_INTERNAL_ENERGY 23
#define DNLG_FIELD_INELASTIC_ENERGY 24
#define DNLG_FIELD_KINETIC_ENERGY 25
#define DNLG_FIELD_MASS 26
#define DNLG_FIELD_PLASTIC_STRAIN 27
#define DNLG_FIELD_PLASTIC_STRAIN_RATE 28
#define DNLG_FIELD_PRESSURE 29
#define DNLG_FIELD_REAL_TIME_STEP 30
#define DNLG_FIELD_T 31
#define DNLG_FIELD_TIME_STEP 32
#define DNLG_FIELD_VON_MISES 33
#define DNLG_FIELD_YIELD_STRESS 34
#define DNLG_FIELD_PLASTIC_STRAIN_TENSOR 35 /* PlasticStrain */
#define DNLG_FIELD_PLASTIC_STRAIN_INC 45
#define DNLG_FIELD_STRAIN 55
#define DNLG_FIELD_STRAIN_INC 65
#define DNLG_FIELD_STRESS 75
#define DNLG_FIELD_COUNT 85
/* item kinds */
#define DNLG_H_NODE 0    /* HistoryFile::add(NodeSet *, field), one item per node:
                            Node::fieldScalar (Node.C:311-370), IP fields
                            extrapolated to the node and averaged over its
                            elements */
#define DNLG_H_ELEMENT 1 /* HistoryFile::add(ElementSet *, intPt, field), one
                            item per element: Element::getIntPointValue
                            (Element.C:1195-1238) */
#define DNLG_H_GLOBAL 2  /* HistoryFile::add(field): timeStep, realTimeStep,
                            kineticEnergy (HistoryFile.C:80-99) */
typedef struct dnlg_history_item {
  int32_t kind;    /* DNLG_H_* */
  int32_t field;   /* DNLG_FIELD_* (+ component offset) */
  int64_t index;   /* node or element index (0-based, mesh order) */
  int32_t int_pt;  /* element items: 1..n_ip, 0 = mean over the IPs */
  int32_t reserved;
} dnlg_history_item;
/* Replaces HistoryFile::setFileName / setSaveTime / add + Model::add
 * (HistoryFile.C:131-280, Model.C:927-939) for a job that has not been
 * solved yet.  path: the file, created now with the reference's header
 * (HistoryFile::headerWrite, HistoryFile.C:340-358); NULL keeps the rows in
 * memory only.  names: the n_items header names (may be NULL).  A field
 * that the reference's getValue does not know for the item's kind gives 0,
 * as there (e.g. inelasticEnergy); StrainInc and PlasticStrainInc are not
 * kept by the engine (DNLG_E_INVALID); Strain and PlasticStrain need
 * track_tensors (DNLG_E_STATE).  *file_id: the file's index. */
int dnlg_job_add_history(dnlg_job *job, const char *path, double start_time,
                         double save_time, const dnlg_history_item *items,
                         int32_t n_items, const char *const *names,
                         int32_t *file_id);
/* The rows the file has taken so far (*n_rows); rows != NULL: the first
 * min(*n_rows, max_rows) rows of 1 + n_items doubles are copied out. */
int dnlg_job_history_rows(dnlg_job *job, int32_t file_id, double *rows,
                          int64_t max_rows, int64_t *n_rows);

/* ---- result (VTK) files and node fields -------------------------------------
 * DynELA::setSaveTimes / writeResultFile / writeVTKFile (DynELA.C:479-503,
 * 688-710) with VtkInterface::write (VtkInterface.C:78-278).  Set before the
 * first solve: base name (files <base><index, 3 digits>.vtk), save times and
 * the point-data fields (DNLG_FIELD_* values: a vector or tensor block's
 * first value writes VECTORS / TENSORS, anything else SCALARS; fields the
 * engine does not keep -- the increments, strains without track_tensors --
 * are left out).  During a solve a file is written after every cycle whose
 * new time t has next <= t <= end (next += increment): the job pauses at
 * that cycle until the host has written the file from the device state, the
 * others go on. */
int dnlg_job_set_result_files(dnlg_job *job, const char *base_name,
                              double start_time, double end_time,
                              double increment, const int32_t *fields,
                              int32_t n_fields);
/* force = 0: DynELA::writeResultFile's rule at the job's current time (the
 * reference calls it before its first solve); force = 1: DynELA::writeVTKFile
 * (after the solve), a file whatever the time. */
int dnlg_job_write_result_file(dnlg_job *job, int32_t force);
/* Node::fieldScalar (Node.C:311-370) of every node, out [n_nodes]: nodal
 * values, or IP values extrapolated to the nodes (integrationPointsToNode)
 * and averaged over each node's elements; 0 for a field the reference's
 * function does not know. */
int dnlg_job_read_node_field(dnlg_job *job, int32_t field, double *out);

/* ---- launch schedules and the autotuner -------------------------------------
 * How the element kernels and the corrector are laid out on the GPU; never
 * what they compute (results are the same bits under every schedule).  The
 * reference splits its element loops into fixed OpenMP chunks
 * (Parallel.C:78-113); here a schedule per element type, precision and job
 * size comes from a table that the autotuner measured on the device. */
typedef struct dnlg_schedule {
  int32_t elem_block;   /* threads per element-kernel block: hex 64 / 128 / 256,
                           axisymmetric quad 128 / 256 / 512; 0 = default */
  int32_t waves;        /* waves per SIMD asked of the compiler: 0 = the
                           variant's default, or its alternative (mixed 4,
                           mixed-rr32 5, FP64 2; FP64's default is 3) */
  int32_t corr_batch;   /* incidence slots the corrector loads together:
                           2, 4, 8 (FP32 forces) or 2, 4 (FP64); 0 = default */
  int32_t graph_cycles; /* cycles per captured HIP graph; 0 = default */
} dnlg_schedule;
/* Whether the kernels are built for this schedule of a job of this element
 * type, precision and hardening law (1 yes, 0 no); the plane-strain quad and
 * the network law run their default schedule only.  Needs no device. */
int dnlg_schedule_valid(int elem_type, int precision, int law, const dnlg_schedule *sched);
/* A rule of the context's table: jobs of this element type and precision
 * with min_elems <= n_elems < max_elems (max_elems <= 0: no bound) run with
 * this schedule; the first matching rule wins; DNLG_E_INVALID for a schedule
 * the kernels are not built for. */
int dnlg_ctx_add_schedule(dnlg_ctx *ctx, int elem_type, int precision, int64_t min_elems, int64_t max_elems,
                          const dnlg_schedule *sched);
int dnlg_ctx_clear_schedules(dnlg_ctx *ctx);
/* The schedule of one job (overrides the rules), and the one it runs with. */
int dnlg_job_set_schedule(dnlg_job *job, const dnlg_schedule *sched);
int dnlg_job_get_schedule(dnlg_job *job, dnlg_schedule *sched);
/* Autotuner: measures the candidate schedules on this job by advancing it
 * (it must be between solves and at least rounds * (number of candidates +
 * 3) * cycles_per_trial cycles short of its end time): each trial runs
 * cycles_per_trial cycles after one untimed graph's worth, trials go
 * round-robin over the candidates `rounds` times, and a candidate's time is
 * its median.  Coordinate-wise: element block and waves, then the corrector
 * batch, then the graph length.  The best schedule is left set on the job
 * and returned; ms_per_cycle[0] its time per cycle (ms), ms_per_cycle[1]
 * the starting schedule's (as measured among the candidates). */
int dnlg_job_autotune(dnlg_job *job, int64_t cycles_per_trial, int rounds, dnlg_schedule *best,
                      double ms_per_cycle[2]);

/* ---- measurement hooks ------------------------------------------------------ */
/* Device time (ms) and launch count of the element kernels, accumulated with
 * HIP events recorded around them on the engine's own stream since the last
 * reset; name: "element".  With graphs on, one cycle in every graph's worth
 * of cycles is sampled.  Timing is off unless enabled. */
int dnlg_ctx_set_timing(dnlg_ctx *ctx, int enable);
int dnlg_ctx_kernel_time(dnlg_ctx *ctx, const char *name, double *total_ms,
                         int64_t *launches);
int dnlg_ctx_reset_timing(dnlg_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* DNLG_H */
