"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the engine.  It takes the same flat
descriptors as the engine (dnlPython ``model.descriptors()``) so both solve
the same model.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_EIGEN = os.path.join(HERE, "_ref", "libref_eigen3x3.so")

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_lib = None

F_PLASTIC_STRAIN, F_PLASTIC_STRAIN_RATE, F_GAMMA, F_GAMMA_CUMULATE, F_YIELD_STRESS, F_TEMPERATURE, \
    F_DENSITY, F_PRESSURE, F_INTERNAL_ENERGY, F_INELASTIC_ENERGY, F_VON_MISES, F_STRESS, F_STRAIN, \
    F_PLASTIC_STRAIN_TENSOR = range(14)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB)
        lib.oracle_job_create.restype = ctypes.c_int
        lib.oracle_job_create.argtypes = [_vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]
        lib.oracle_job_solve.argtypes = [_vp, ctypes.c_double]
        lib.oracle_job_destroy.argtypes = [_vp]
        lib.oracle_job_set_threads.argtypes = [_vp, ctypes.c_int]
        lib.oracle_job_read_nodes.argtypes = [_vp, _dp, _dp, _dp]
        lib.oracle_job_read_mass.argtypes = [_vp, _dp]
        lib.oracle_job_read_ip.argtypes = [_vp, ctypes.c_int, _dp]
        lib.oracle_job_scalars.argtypes = [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_job_energies.argtypes = [_vp, _dp]
        lib.oracle_job_sizes.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        lib.oracle_job_status.argtypes = [_vp]
        lib.oracle_eig3.argtypes = [_dp, _dp, _dp]
        lib.oracle_polar_lnU.argtypes = [_dp, _dp, _dp]
        lib.oracle_jacobi_eig3.argtypes = [_dp, _dp, _dp]
        lib.oracle_rotate_sym.argtypes = [_dp, _dp]
        lib.oracle_law_yield.restype = ctypes.c_double
        lib.oracle_law_yield.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        lib.oracle_law_dyield.restype = ctypes.c_double
        lib.oracle_law_dyield.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        lib.oracle_radial_return.argtypes = [_vp, _dp, ctypes.c_double, ctypes.c_double, _dp]
        lib.oracle_tables.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        lib.oracle_char_length.restype = ctypes.c_double
        lib.oracle_char_length.argtypes = [ctypes.c_int, _dp]
        lib.oracle_chung_hulbert.argtypes = [ctypes.c_double, _dp]
        lib.oracle_ramp_value.restype = ctypes.c_double
        lib.oracle_ramp_value.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        lib.oracle_element_stiffness.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                 _dp]
        lib.oracle_job_pi_state.argtypes = [_vp, _dp, ctypes.POINTER(ctypes.c_int)]
        lib.oracle_job_add_history.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(HistoryItem),
                                               ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        lib.oracle_job_history_rows.argtypes = [_vp, ctypes.c_int32, _dp, ctypes.c_int64,
                                                ctypes.POINTER(ctypes.c_int64)]
        lib.oracle_extrapolation.argtypes = [ctypes.c_int, _dp]
        lib.oracle_job_read_node_field.argtypes = [_vp, ctypes.c_int, _dp]
        lib.oracle_job_read_state.argtypes = [_vp, _vp]
        lib.oracle_job_write_state.argtypes = [_vp, _vp]
        _lib = lib
    return _lib


def p(a):
    return np.ascontiguousarray(a).ctypes.data_as(_dp)


class HistoryItem(ctypes.Structure):
    """dnlg_history_item of include/dnlg.h."""
    _fields_ = [("kind", ctypes.c_int32), ("field", ctypes.c_int32), ("index", ctypes.c_int64),
                ("int_pt", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def extrapolation(elem_type):
    """The oracle's IP-to-node extrapolation table, [node][ip]."""
    nn = 8 if elem_type == 4 else 4
    nip = 1 if elem_type == 5 else nn  # DNLG_EL_TET4N_3D = 5: one integration point
    E = np.zeros((nn, nip))
    load().oracle_extrapolation(elem_type, E.ctypes.data_as(_dp))
    return E


class OracleJob:
    """CPU FP64 solve of a model descriptor set (dnlPython ModelArrays)."""

    def __init__(self, arrays, threads=1):
        lib = load()
        self.lib = lib
        self.arrays = arrays
        self.h = _vp()
        st = lib.oracle_job_create(ctypes.addressof(arrays.mesh), ctypes.addressof(arrays.material),
                                   ctypes.addressof(arrays.bcs), ctypes.addressof(arrays.params),
                                   ctypes.byref(self.h))
        if st != 0:
            raise RuntimeError(f"oracle_job_create failed: {st}")
        lib.oracle_job_set_threads(self.h, threads)
        n_nodes, n_elems = ctypes.c_int64(), ctypes.c_int64()
        nip, dim = ctypes.c_int32(), ctypes.c_int32()
        lib.oracle_job_sizes(self.h, ctypes.byref(n_nodes), ctypes.byref(n_elems), ctypes.byref(nip),
                             ctypes.byref(dim))
        self.N, self.NE, self.nip, self.dim = n_nodes.value, n_elems.value, nip.value, dim.value
        self._hist_width = {}

    def solve(self, up_to):
        return self.lib.oracle_job_solve(self.h, ctypes.c_double(up_to))

    def nodes(self):
        c, d, v = np.zeros((3, self.N)), np.zeros((3, self.N)), np.zeros((3, self.N))
        self.lib.oracle_job_read_nodes(self.h, p(c), p(d), p(v))
        return c, d, v

    def mass(self):
        m = np.zeros(self.N)
        self.lib.oracle_job_read_mass(self.h, p(m))
        return m

    def ip(self, field):
        ncomp = 6 if field in (F_STRESS, F_STRAIN, F_PLASTIC_STRAIN_TENSOR) else 1
        out = np.zeros(ncomp * self.NE * self.nip)
        st = self.lib.oracle_job_read_ip(self.h, field, p(out))
        if st != 0:
            raise RuntimeError(f"oracle_job_read_ip failed: {st}")
        return out.reshape((ncomp, self.NE, self.nip) if ncomp == 6 else (self.NE, self.nip))

    def scalars(self):
        t, dt, inc = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self.lib.oracle_job_scalars(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(inc))
        return t.value, dt.value, inc.value

    def energies(self):
        out = np.zeros(4)
        self.lib.oracle_job_energies(self.h, p(out))
        return out

    def status(self):
        return self.lib.oracle_job_status(self.h)

    def add_history(self, start, save, items):
        """A history file with items (kind, field, index, int_pt); returns its id."""
        n = len(items)
        arr = (HistoryItem * max(1, n))(*[HistoryItem(k, f, i, ip, 0) for (k, f, i, ip) in items])
        fid = ctypes.c_int32()
        st = self.lib.oracle_job_add_history(self.h, start, save, arr, n, ctypes.byref(fid))
        if st != 0:
            raise RuntimeError(f"oracle_job_add_history failed: {st}")
        self._hist_width[fid.value] = 1 + n
        return fid.value

    def add_history_file(self, model, hist):
        """The items of a dnlPython HistoryFile attached to `model`, numbers
        resolved to the model's node / element indices."""
        model.descriptors()
        items = []
        for kind, f, num, ip, _ in hist._items:
            if kind == 0:
                idx = int(np.searchsorted(model._node_sorted, num))
            elif kind == 1:
                idx = int(np.searchsorted(model._elem_sorted, num))
            else:
                idx = 0
            items.append((kind, f, idx, ip))
        return self.add_history(hist.getStartTime(), hist.getSaveTime(), items)

    def node_field(self, field):
        """Node::fieldScalar of every node (Field enum value)."""
        out = np.zeros(self.N)
        st = self.lib.oracle_job_read_node_field(self.h, int(field), p(out))
        if st != 0:
            raise RuntimeError(f"oracle_job_read_node_field failed: {st}")
        return out

    def history_rows(self, fid):
        """Rows taken so far, [rows, 1 + items] (time first)."""
        n = ctypes.c_int64()
        self.lib.oracle_job_history_rows(self.h, fid, None, 0, ctypes.byref(n))
        out = np.zeros((n.value, self._hist_width[fid]))
        if n.value:
            self.lib.oracle_job_history_rows(self.h, fid, p(out), n.value, ctypes.byref(n))
        return out

    def get_state(self):
        """The solver state (oracle_job_read_state), as dnlPython's
        DynELA.getState returns it."""
        from dnlPython import _lib as dl
        arr, rec = dl.state_alloc(self.N, self.NE * self.nip, True)
        self.lib.oracle_job_read_state(self.h, ctypes.addressof(rec))
        arr.update(time=rec.time, timeStep=rec.time_step, increment=rec.increment)
        return arr

    def set_state(self, state):
        """Continue from a state of DynELA.getState or get_state
        (oracle_job_write_state); returns the status."""
        from dnlPython import _lib as dl
        arr = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in state.items() if isinstance(v, np.ndarray)}
        rec = dl.state_record(arr, float(state["time"]), float(state["timeStep"]), int(state["increment"]))
        return self.lib.oracle_job_write_state(self.h, ctypes.addressof(rec))

    def pi_state(self):
        """(last power-iteration maximum, iterations of the last call)"""
        f, it = ctypes.c_double(), ctypes.c_int()
        self.lib.oracle_job_pi_state(self.h, ctypes.byref(f), ctypes.byref(it))
        return f.value, it.value

    def __del__(self):
        try:
            if self.h:
                self.lib.oracle_job_destroy(self.h)
        except Exception:
            pass


def element_stiffness(elem_type, xyz, young, poisson, under=False):
    """Elastic stiffness matrix (nn dim x nn dim) of one element with node
    coordinates xyz[a] = (x, y, z); under: the one-point under-integrated
    matrix (PowerIterationUnder)."""
    lib = load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    nd = xyz.shape[0] * (3 if elem_type in (4, 5) else 2)  # DNLG_EL_HEX8N_3D = 4, DNLG_EL_TET4N_3D = 5
    K = np.zeros((nd, nd))
    st = lib.oracle_element_stiffness(elem_type, p(xyz), young, poisson, 1 if under else 0, K.ctypes.data_as(_dp))
    if st != 0:
        raise RuntimeError(f"oracle_element_stiffness failed: {st}")
    return K
