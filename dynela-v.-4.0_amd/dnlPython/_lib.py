"""ctypes binding of the engine's C ABI (include/dnlg.h, libdnlg.so).

The structures below mirror dnlg.h field by field.  The library is loaded
from the package directory; there is no CPU fallback: if libdnlg.so is
missing or no GPU is visible, creating a context raises.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libdnlg.so")

# ---- constants of dnlg.h ---------------------------------------------------
OK = 0
E_INVALID, E_OOM, E_NEG_JACOBIAN, E_NR_DIVERGED, E_NONFINITE, E_HIP, E_STATE, E_CAPACITY, E_PI_DIVERGED = (
    -1, -2, -3, -4, -5, -6, -7, -8, -9)
E_TABLE_RANGE = -10
STATUS_NAMES = {
    OK: "OK", E_INVALID: "INVALID", E_OOM: "OUT_OF_MEMORY", E_NEG_JACOBIAN: "NEGATIVE_JACOBIAN",
    E_NR_DIVERGED: "NEWTON_RAPHSON_DIVERGED", E_NONFINITE: "NON_FINITE", E_HIP: "HIP_ERROR",
    E_STATE: "INVALID_STATE", E_CAPACITY: "CAPACITY", E_PI_DIVERGED: "POWER_ITERATION_DIVERGED",
    E_TABLE_RANGE: "TABLE_RANGE",
}
EL_QUA4N_2D = 1
EL_QUA4N_AX = 3
EL_HEX8N_3D = 4
EL_TET4N_3D = 5
LAW_JOHNSON_COOK, LAW_POWER, LAW_VOCE_LINEAR, LAW_TABULAR, LAW_ANN2SIG = 0, 1, 2, 3, 4
ANN_MAX_HIDDEN = 64
BC_RESTRAIN, BC_SPEED = 0, 1
RAMP_NONE, RAMP_LINEAR, RAMP_SINUS, RAMP_CONSTANT, RAMP_ERF, RAMP_SOFT = range(6)
DT_COURANT, DT_POWER_ITERATION, DT_POWER_ITERATION_UNDER = 0, 1, 2
STRESS_NR, STRESS_DIRECT = 0, 1
PREC_FP64, PREC_MIXED, PREC_MIXED_RR32 = 0, 1, 2
(F_PLASTIC_STRAIN, F_PLASTIC_STRAIN_RATE, F_GAMMA, F_GAMMA_CUMULATE, F_YIELD_STRESS, F_TEMPERATURE,
 F_DENSITY, F_PRESSURE, F_INTERNAL_ENERGY, F_INELASTIC_ENERGY, F_VON_MISES, F_STRESS, F_STRAIN,
 F_PLASTIC_STRAIN_TENSOR) = range(14)
TENSOR_FIELDS = (F_STRESS, F_STRAIN, F_PLASTIC_STRAIN_TENSOR)
H_NODE, H_ELEMENT, H_GLOBAL = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class Mesh(ctypes.Structure):
    _fields_ = [("elem_type", ctypes.c_int32), ("reserved", ctypes.c_int32), ("n_nodes", ctypes.c_int64),
                ("n_elems", ctypes.c_int64), ("coords", _dp), ("conn", _i32p)]


class MaterialDesc(ctypes.Structure):
    _fields_ = [("young", ctypes.c_double), ("poisson", ctypes.c_double), ("density", ctypes.c_double),
                ("heat_capacity", ctypes.c_double), ("taylor_quinney", ctypes.c_double),
                ("T0", ctypes.c_double), ("law", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("p", ctypes.c_double * 8), ("table", ctypes.POINTER(ctypes.c_double)), ("n_table", ctypes.c_int64)]


class BoundaryDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ramp", ctypes.c_int32), ("value", ctypes.c_double * 3),
                ("ramp_min", ctypes.c_double), ("ramp_max", ctypes.c_double)]


class Bcs(ctypes.Structure):
    _fields_ = [("n_boundaries", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("boundaries", ctypes.POINTER(BoundaryDesc)), ("initial_ptr", _i64p),
                ("initial_idx", _i32p), ("constant_ptr", _i64p), ("constant_idx", _i32p)]


class State(ctypes.Structure):
    """dnlg_state of include/dnlg.h (checkpoint / restart record)."""
    _fields_ = [("time", ctypes.c_double), ("time_step", ctypes.c_double), ("increment", ctypes.c_int64),
                ("coords", _dp), ("disp_inc", _dp), ("speed", _dp), ("accel", _dp), ("stress", _dp),
                ("ip_scalars", _dp), ("strain", _dp), ("plastic_strain_tensor", _dp)]


STATE_IP_NAMES = ("plasticStrain", "plasticStrainRate", "gamma", "gammaCumulate", "yieldStress", "T",
                    "internalEnergy", "inelasticEnergy", "density")
STATE_NODAL = ("coords", "disp_inc", "speed", "accel")
STATE_IP_SCALARS = len(STATE_IP_NAMES)


def state_alloc(n_nodes, n_ip_total, tensors):
    """numpy arrays of a state record of these sizes, and the record
    pointing at them."""
    arr = {k: np.zeros((3, n_nodes)) for k in STATE_NODAL}
    arr["stress"] = np.zeros((6, n_ip_total))
    arr["ip_scalars"] = np.zeros((len(STATE_IP_NAMES), n_ip_total))
    if tensors:
        arr["strain"] = np.zeros((6, n_ip_total))
        arr["plastic_strain_tensor"] = np.zeros((6, n_ip_total))
    return arr, state_record(arr)


def state_record(arr, time=0.0, time_step=0.0, increment=0):
    """A State record pointing at the (C-contiguous, float64) arrays."""
    s = State()
    s.time, s.time_step, s.increment = time, time_step, increment
    for k in STATE_NODAL + ("stress", "ip_scalars", "strain", "plastic_strain_tensor"):
        a = arr.get(k)
        if a is not None:
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
            setattr(s, k, a.ctypes.data_as(_dp))
    return s


class Schedule(ctypes.Structure):
    """dnlg_schedule of include/dnlg.h (launch schedule of a job)."""
    _fields_ = [("elem_block", ctypes.c_int32), ("waves", ctypes.c_int32), ("corr_batch", ctypes.c_int32),
                ("graph_cycles", ctypes.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class SolverParams(ctypes.Structure):
    _fields_ = [("t_start", ctypes.c_double), ("t_end", ctypes.c_double), ("rho_b", ctypes.c_double),
                ("safety", ctypes.c_double), ("dt_method", ctypes.c_int32), ("dt_freq", ctypes.c_int32),
                ("stress_method", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("track_tensors", ctypes.c_int32), ("nr_max_iter", ctypes.c_int32)]


class HistoryItem(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("field", ctypes.c_int32), ("index", ctypes.c_int64),
                ("int_pt", ctypes.c_int32), ("reserved", ctypes.c_int32)]


# every entry point of include/dnlg.h: name -> (restype, argtypes)
_VP = ctypes.c_void_p
SIGNATURES = {
    "dnlg_abi_version": (ctypes.c_int, []),
    "dnlg_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_VP)]),
    "dnlg_ctx_destroy": (ctypes.c_int, [_VP]),
    "dnlg_last_error": (ctypes.c_char_p, [_VP]),
    "dnlg_job_create": (ctypes.c_int, [_VP, ctypes.POINTER(Mesh), ctypes.POINTER(MaterialDesc),
                                       ctypes.POINTER(Bcs), ctypes.POINTER(SolverParams),
                                       ctypes.POINTER(_VP)]),
    "dnlg_job_destroy": (ctypes.c_int, [_VP]),
    "dnlg_job_solve": (ctypes.c_int, [_VP, ctypes.c_double]),
    "dnlg_batch_solve": (ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.c_int, _dp]),
    "dnlg_batch_advance": (ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.c_int, _dp, ctypes.c_int64]),
    "dnlg_job_sizes": (ctypes.c_int, [_VP, _i64p, _i64p, _i32p, _i32p]),
    "dnlg_job_read_nodes": (ctypes.c_int, [_VP, _dp, _dp, _dp]),
    "dnlg_job_read_ip": (ctypes.c_int, [_VP, ctypes.c_int, _dp]),
    "dnlg_job_scalars": (ctypes.c_int, [_VP, _dp, _dp, _i64p]),
    "dnlg_job_energies": (ctypes.c_int, [_VP, _dp]),
    "dnlg_job_energies_static": (ctypes.c_int, [_VP, _dp]),
    "dnlg_job_read_state": (ctypes.c_int, [_VP, ctypes.POINTER(State)]),
    "dnlg_schedule_valid": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Schedule)]),
    "dnlg_ctx_add_schedule": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.POINTER(Schedule)]),
    "dnlg_ctx_clear_schedules": (ctypes.c_int, [_VP]),
    "dnlg_job_set_schedule": (ctypes.c_int, [_VP, ctypes.POINTER(Schedule)]),
    "dnlg_job_get_schedule": (ctypes.c_int, [_VP, ctypes.POINTER(Schedule)]),
    "dnlg_job_autotune": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(Schedule), _dp]),
    "dnlg_job_write_state": (ctypes.c_int, [_VP, ctypes.POINTER(State)]),
    "dnlg_job_read_mass": (ctypes.c_int, [_VP, _dp]),
    "dnlg_footprint_bytes": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                              ctypes.c_int]),
    "dnlg_model_footprint_bytes": (ctypes.c_int64, [_VP, _VP, _VP, _VP]),
    "dnlg_job_status": (ctypes.c_int, [_VP]),
    "dnlg_ctx_set_timing": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dnlg_ctx_kernel_time": (ctypes.c_int, [_VP, ctypes.c_char_p, _dp, _i64p]),
    "dnlg_ctx_reset_timing": (ctypes.c_int, [_VP]),
    "dnlg_server_create": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(_VP)]),
    "dnlg_server_destroy": (ctypes.c_int, [_VP]),
    "dnlg_server_submit": (ctypes.c_int, [_VP, ctypes.POINTER(Mesh), ctypes.POINTER(MaterialDesc),
                                          ctypes.POINTER(Bcs), ctypes.POINTER(SolverParams), _i64p]),
    "dnlg_server_step": (ctypes.c_int, [_VP, ctypes.c_int, _i32p, _i32p]),
    "dnlg_server_set_cycle_budget": (ctypes.c_int, [_VP, ctypes.c_int64]),
    "dnlg_server_poll": (ctypes.c_int, [_VP, ctypes.c_int64, _i32p, _i32p]),
    "dnlg_server_take": (ctypes.c_int, [_VP, ctypes.c_int64, ctypes.POINTER(_VP)]),
    "dnlg_job_add_history": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                            ctypes.POINTER(HistoryItem), ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_char_p), _i32p]),
    "dnlg_job_history_rows": (ctypes.c_int, [_VP, ctypes.c_int32, _dp, ctypes.c_int64, _i64p]),
    "dnlg_job_set_result_files": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_double, _i32p, ctypes.c_int32]),
    "dnlg_job_write_result_file": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "dnlg_job_read_node_field": (ctypes.c_int, [_VP, ctypes.c_int32, _dp]),
}

_lib = None


def load(path=None):
    """Load libdnlg.so (once) and declare every signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libdnlg.so not found at {p}: build it with `python build.py` in {PKG_DIR}")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.dnlg_abi_version() != 2:
        raise RuntimeError("libdnlg.so ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


class DnlgError(RuntimeError):
    """Error status returned by the engine."""

    def __init__(self, status, message=""):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


class NegativeJacobianError(DnlgError):
    pass


class NewtonRaphsonError(DnlgError):
    pass


class NonFiniteError(DnlgError):
    pass


class PowerIterationError(DnlgError):
    pass


_ERRORS = {E_NEG_JACOBIAN: NegativeJacobianError, E_NR_DIVERGED: NewtonRaphsonError,
           E_NONFINITE: NonFiniteError, E_PI_DIVERGED: PowerIterationError}


def check(status, ctx=None):
    if status != OK:
        msg = ""
        if ctx is not None and _lib is not None:
            m = _lib.dnlg_last_error(ctx)
            msg = m.decode() if m else ""
        raise _ERRORS.get(status, DnlgError)(status, msg)
    return status


def as_dp(a):
    return a.ctypes.data_as(_dp)


def as_i32p(a):
    return a.ctypes.data_as(_i32p)


def as_i64p(a):
    return a.ctypes.data_as(_i64p)


class ModelArrays:
    """Flat descriptor arrays of one model, kept alive with their ctypes
    views (the engine copies them at job creation)."""

    def __init__(self, elem_type, coords, conn, material, boundaries, initial_lists, constant_lists,
                 params):
        self.elem_type = int(elem_type)
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)  # [3, N]
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)        # [NE, nn]
        n_nodes = self.coords.shape[1]
        self.mesh = Mesh(self.elem_type, 0, n_nodes, self.conn.shape[0], as_dp(self.coords),
                         as_i32p(self.conn))
        self.material = material
        self.boundaries = (BoundaryDesc * max(1, len(boundaries)))(*boundaries)

        def csr(lists):
            """(ptr [N+1] int64, idx int32) or None."""
            if lists is None:
                return None, None
            ptr, idx = lists
            ptr = np.ascontiguousarray(ptr, dtype=np.int64)
            idx = np.ascontiguousarray(idx if len(idx) else [0], dtype=np.int32)
            if ptr.shape != (n_nodes + 1,):
                raise ValueError("boundary list pointer has the wrong length")
            return ptr, idx

        self.init_ptr, self.init_idx = csr(initial_lists)
        self.const_ptr, self.const_idx = csr(constant_lists)
        self.bcs = Bcs(len(boundaries), 0, self.boundaries,
                       as_i64p(self.init_ptr) if self.init_ptr is not None else None,
                       as_i32p(self.init_idx) if self.init_idx is not None else None,
                       as_i64p(self.const_ptr) if self.const_ptr is not None else None,
                       as_i32p(self.const_idx) if self.const_idx is not None else None)
        self.params = params
