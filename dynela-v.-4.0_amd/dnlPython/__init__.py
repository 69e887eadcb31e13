"""dnlPython -- drop-in model-building API of DynELA's explicit solve, run on
the MI355X engine (libdnlg.so).

Mirrors the class and method names of the reference's SWIG module
(Sources/dnlPython/dnlPython.i:10-25, dnlFEM.inc:11-35): ``DynELA``,
``NodeSet``, ``ElementSet``, ``Material``, ``JohnsonCookLaw``,
``ElastoplasticLaw``, ``BoundaryRestrain``, ``BoundarySpeed``,
``RampFunction``, ``Explicit``, ``Field``, ``Element``.  ``model.solve()``
replaces ``Explicit::solve`` (Sources/dnlFEM/Explicit.C:126-253) with the
GPU engine.  Unlike the reference (one global DynELA object,
Sources/dnlFEM/DynELA.C:21,47-52) any number of models may coexist, and bulk
numpy ingestion (``createNodes`` / ``createElements``) replaces the O(N^2)
per-entity build path (Model.C:91,122) for large meshes.

Errors never exit the interpreter: engine statuses become exceptions
(``NegativeJacobianError``, ``NewtonRaphsonError``, ``NonFiniteError``,
``PowerIterationError``, ``DnlgError``).
"""
import ctypes
import os
import time
import warnings
import weakref

import numpy as np

from . import _lib
from ._lib import (DnlgError, NegativeJacobianError, NewtonRaphsonError, PowerIterationError,  # noqa: F401
                   NonFiniteError)

__all__ = ["DynELA", "NodeSet", "ElementSet", "Material", "JohnsonCookLaw", "ElastoplasticLaw", "HistoryFile",
           "TabularLaw", "DiscreteFunction", "ANN2SigLaw", "ANNFortranLaw", "ElasticLaw",
           "VtkInterface", "SvgInterface",
           "VoceLinearLaw", "BoundaryRestrain", "BoundarySpeed", "RampFunction", "Explicit", "Field",
           "Element", "Vec3D", "Context", "DnlgError", "NegativeJacobianError", "NewtonRaphsonError",
           "PowerIterationError", "NonFiniteError", "batch_solve"]


class Element:
    """Element type names of the reference enum (Element.h:62-71)."""
    ElGeneric = 0
    ElQua4N2D = 1
    ElTri3N2D = 2
    ElQua4NAx = _lib.EL_QUA4N_AX
    ElHex8N3D = _lib.EL_HEX8N_3D
    ElTet4N3D = _lib.EL_TET4N_3D
    ElTet10N3D = 6


NODES_PER_ELEMENT = {Element.ElQua4N2D: 4, Element.ElQua4NAx: 4, Element.ElHex8N3D: 8, Element.ElTet4N3D: 4}
# integration points per element (ElTet4N3D.C:62-71: one point)
IPS_PER_ELEMENT = {Element.ElQua4N2D: 4, Element.ElQua4NAx: 4, Element.ElHex8N3D: 8, Element.ElTet4N3D: 1}


def _field_names():
    """The identifiers of Field::FieldLabel in enum order (Field.h:46-79):
    the values are the enum's, as the reference's Python module exposes them."""
    names = []
    for v in ("coords", "disp", "dispInc", "speed", "speedInc"):
        names += [v, v + "X", v + "Y", v + "Z"]
    names += ["density", "gamma", "gammaCumulate", "internalEnergy", "inelasticEnergy", "kineticEnergy", "mass",
              "plasticStrain", "plasticStrainRate", "pressure", "realTimeStep", "T", "timeStep", "vonMises",
              "yieldStress"]
    for t in ("PlasticStrain", "PlasticStrainInc", "Strain", "StrainInc", "Stress"):
        names += [t] + [t + c for c in ("XX", "XY", "XZ", "YX", "YY", "YZ", "ZX", "ZY", "ZZ")]
    return names


_FIELD_NAMES = _field_names()


class Field:
    """Field labels of Field.h:46-79 with the reference's enum values."""

    @staticmethod
    def getVtklabel(field):
        """Field::getVtklabel (Field.C): the identifier, except yieldStress,
        whose label is "yield"."""
        name = _FIELD_NAMES[int(field)]
        return "yield" if name == "yieldStress" else name

    @staticmethod
    def getField(name):
        """Field::getField: the value of a label, -1 if unknown."""
        for f in range(len(_FIELD_NAMES)):
            if Field.getVtklabel(f) == name:
                return f
        return -1


for _i, _n in enumerate(_FIELD_NAMES):
    setattr(Field, _n, _i)
Field.ENDFIELDS = len(_FIELD_NAMES)


_IP_FIELDS = {
    Field.plasticStrain: _lib.F_PLASTIC_STRAIN, Field.plasticStrainRate: _lib.F_PLASTIC_STRAIN_RATE,
    Field.gamma: _lib.F_GAMMA, Field.gammaCumulate: _lib.F_GAMMA_CUMULATE,
    Field.yieldStress: _lib.F_YIELD_STRESS, Field.T: _lib.F_TEMPERATURE, Field.density: _lib.F_DENSITY,
    Field.pressure: _lib.F_PRESSURE, Field.internalEnergy: _lib.F_INTERNAL_ENERGY,
    Field.inelasticEnergy: _lib.F_INELASTIC_ENERGY, Field.vonMises: _lib.F_VON_MISES,
    Field.Stress: _lib.F_STRESS, Field.Strain: _lib.F_STRAIN,
    Field.PlasticStrain: _lib.F_PLASTIC_STRAIN_TENSOR,
}


class Vec3D:
    def __init__(self, x=0.0, y=0.0, z=0.0):
        self._v = [float(x), float(y), float(z)]

    def __call__(self, i):
        return self._v[i]

    def __getitem__(self, i):
        return self._v[i]

    def __repr__(self):
        return f"Vec3D({self._v[0]}, {self._v[1]}, {self._v[2]})"


class _MemberSet:
    """Members of a NodeSet / ElementSet by number, in insertion order, with
    the reference's set operations (NodeSet.C:87-160, ElementSet.C)."""

    def __init__(self, name=""):
        self.name = name
        self._numbers = []

    def setName(self, name):
        self.name = name

    def size(self):
        return len(self._numbers)

    def getNumber(self, item):
        return self._numbers[item]

    def clear(self):
        self._numbers = []

    def intersect(self, other):
        """Keeps the members that are also in `other`."""
        keep = set(other._numbers)
        self._numbers = [n for n in self._numbers if n in keep]

    def substract(self, other):
        """Removes the members that are in `other`."""
        drop = set(other._numbers)
        self._numbers = [n for n in self._numbers if n not in drop]


class NodeSet(_MemberSet):
    def numbers(self):
        return np.array(self._numbers, dtype=np.int64)


class ElementSet(_MemberSet):
    pass


class JohnsonCookLaw:
    """JohnsonCookLaw::setParameters(A, B, C, n, m, depsp0, Tm, T0)
    (JohnsonCookLaw.h:77)."""
    law = _lib.LAW_JOHNSON_COOK

    def __init__(self, name=""):
        self.name = name
        self.p = [0.0] * 8

    def setParameters(self, A, B, C, n, m, depsp0, Tm, T0):
        self.p = [float(A), float(B), float(C), float(n), float(m), float(depsp0), float(Tm), float(T0)]


class ElastoplasticLaw:
    """ElastoplasticLaw::setParameters(A, B, n) (ElastoplasticLaw.h:46)."""
    law = _lib.LAW_POWER

    def __init__(self, name=""):
        self.name = name
        self.p = [0.0] * 8

    def setParameters(self, A, B, n):
        self.p = [float(A), float(B), float(n)] + [0.0] * 5


class ElasticLaw:
    """ElasticLaw (ElasticLaw.C:67-78): yield stress and slope 0.  The
    explicit solve's return (Element.C:470-480) does not distinguish it
    from the other laws, so a material with this law flows at zero
    deviatoric stress: the power law with A = B = 0."""
    law = _lib.LAW_POWER

    def __init__(self, name=""):
        self.name = name
        self.p = [0.0, 0.0, 1.0] + [0.0] * 5

    def setParameters(self, *args):
        pass


class VoceLinearLaw:
    """sigma_y = A + Q (1 - exp(-b ep)) + H ep; no reference counterpart."""
    law = _lib.LAW_VOCE_LINEAR

    def __init__(self, name=""):
        self.name = name
        self.p = [0.0] * 8

    def setParameters(self, A, Q, b, H):
        self.p = [float(A), float(Q), float(b), float(H)] + [0.0] * 4


class DiscreteFunction:
    """DiscreteFunction (DiscreteFunction.h): points (x, y) kept in
    ascending x; getValue / getSlope as DiscreteFunction.C:188-216,
    284-308 (a value outside [x_first, x_last] raises, where the reference
    stops)."""

    def __init__(self, name=""):
        self.name = name
        self._pts = []

    def add(self, x, y):
        self._pts.append((float(x), float(y)))
        self._pts.sort(key=lambda p: p[0])

    def getNumberOfPoints(self):
        return len(self._pts)

    def getXPoint(self, i):
        return self._pts[i][0]

    def getYPoint(self, i):
        return self._pts[i][1]

    def _segment(self, x):
        if not self._pts or x < self._pts[0][0] or x > self._pts[-1][0]:
            raise DnlgError(_lib.E_TABLE_RANGE, f"value {x} is out of bounds")
        for i in range(1, len(self._pts)):
            if self._pts[i][0] >= x:
                return self._pts[i - 1], self._pts[i]
        return None

    def getValue(self, x):
        seg = self._segment(x)
        if seg is None:
            return 0.0
        (x0, y0), (x1, y1) = seg
        return y1 if x1 == x else ((x - x0) / (x1 - x0)) * (y1 - y0) + y0

    def getSlope(self, x):
        seg = self._segment(x)
        if seg is None:
            return 0.0
        (x0, y0), (x1, y1) = seg
        return (y1 - y0) / (x1 - x0)


class TabularLaw:
    """TabularLaw (TabularLaw.C): yield stress as a DiscreteFunction of
    the equivalent plastic strain (FP64 and mixed precision; the mixed path
    evaluates it in its FP64 radial return; not "mixed-rr32")."""
    law = _lib.LAW_TABULAR

    def __init__(self, name=""):
        self.name = name
        self.p = [0.0] * 8
        self._function = None

    def setFunction(self, function):
        self._function = function

    def getFunction(self):
        return self._function


class ANN2SigLaw:
    """ANN2SigLaw (ANN2SigLaw.C): yield stress from a network of 3 inputs
    (plastic strain, log plastic strain rate, temperature), two sigmoid
    hidden layers and one linear output, read by setParameters(filename)
    from a NumPy .npz file with the arrays w1 (3, h1), b1 (h1, 1),
    w2 (h1, h2), b2 (h2, 1), w3 (h2, 1), b3 (1, 1), logBase,
    minEntries (4) and maxEntries (4) (ANN2SigLaw.C:137-158: the weight
    matrices are transposed on reading, the biases are their first
    column).  FP64 and mixed precision (the mixed path evaluates it in its
    FP64 radial return); not "mixed-rr32"."""
    law = _lib.LAW_ANN2SIG
    _NAMES = ("w1", "b1", "w2", "b2", "w3", "b3", "logBase", "minEntries", "maxEntries")

    def __init__(self, name=""):
        self.name = name
        self.p = [0.0] * 8
        self._table = None

    def getNumberOfParameters(self):
        return len(self._NAMES)

    def getParameterName(self, parameter):
        return self._NAMES[parameter]

    def setParameters(self, filename):
        with np.load(str(filename), allow_pickle=False) as z:
            missing = [k for k in self._NAMES if k not in z.files]
            if missing:
                raise DnlgError(_lib.E_INVALID, f"{filename}: no array {missing[0]!r}")
            arrays = {k: np.asarray(z[k], dtype=np.float64) for k in self._NAMES}

        def matrix(k):
            a = arrays[k]
            if a.ndim != 2:
                raise DnlgError(_lib.E_INVALID, f"{filename}: {k} must be a 2-D array")
            return a.T

        def column(k):
            a = arrays[k]
            return a[:, 0] if a.ndim == 2 else a.reshape(-1)

        w1, w2, w3 = matrix("w1"), matrix("w2"), matrix("w3")
        b1, b2, b3 = column("b1"), column("b2"), column("b3")
        h1, h2 = w1.shape[0], w2.shape[0]
        if (w1.shape[1] != 3 or w2.shape[1] != h1 or w3.shape[1] != h2 or len(b1) != h1 or len(b2) != h2
                or len(b3) < 1 or not 1 <= h1 <= _lib.ANN_MAX_HIDDEN or not 1 <= h2 <= _lib.ANN_MAX_HIDDEN):
            raise DnlgError(_lib.E_INVALID, f"{filename}: inconsistent network shapes")
        base = arrays["logBase"].reshape(-1)
        mn, mx = arrays["minEntries"].reshape(-1), arrays["maxEntries"].reshape(-1)
        if len(base) < 1 or len(mn) < 4 or len(mx) < 4:
            raise DnlgError(_lib.E_INVALID, f"{filename}: logBase, minEntries or maxEntries too short")
        self._table = np.concatenate([[h1, h2], w1.ravel(), b1, w2.ravel(), b2, w3[0], b3[:1], base[:1], mn[:4],
                                      mx[:4]]).astype(np.float64)


class ANNFortranLaw(ANN2SigLaw):
    """ANNFortranLaw (ANNFortranLaw.C:76-114, 180-266): the fixed 3-7-4-1
    sigmoid network written out in the reference's source (yield stress in
    MPa).  Its formulas are ANN2SigLaw's for this network, so it is that law
    with these weights.  The constants below keep the reference's form:
    layer 1 is 1 + exp(A x + a), layer 2 exp(C / zb + c), the output
    scale * (E / (1 + zc) + e) + offset, inputs (ep, log(rate) / 10.8198,
    (T - 20) / 480); ANN2SigLaw's exp(-(w x + b)) takes w = -A, b = -a
    and w = -C, b = -c.  setParameters ignores its argument, as the
    reference's does."""
    _A = ((-0.119231261313, -0.173828095198, 1.827572464943), (-0.428523629904, -0.572129368782, 1.296182870865),
          (132.502944946289, -0.074804469943, 0.372308045626), (-0.238687828183, -0.654297590256, -1.221466302872),
          (-0.216847836971, -0.142032980919, 3.807989597321), (16.061351776123, -0.11558868736, 0.41205266118),
          (3.10570526123, -0.138235300779, 0.773996531963))
    _a = (-1.533527970314, 1.823030114174, 1.611178278923, -0.61216711998, -0.042427442968, 1.614393234253,
          1.072145581245)
    _C = ((1.296875834465, 1.162090778351, 10.487031936646, 0.098825834692, 0.730871379375, -4.092911243439,
           1.342621445656),
          (-0.416251301765, 1.612317681313, -1.245551109314, 0.976023316383, 0.005203731358, -1.697751402855,
           -0.834678351879),
          (-1.942471027374, -0.364910364151, 10.365681648254, -0.165790036321, -0.691674530506, -5.159326076508,
           2.959228277206),
          (-1.463814258575, -0.175867184997, 0.67038166523, -0.306437820196, -0.191420003772, 2.523961305618,
           0.034361626953))
    _c = (0.492944657803, -0.692925989628, -0.345911204815, 0.274670749903)
    _E = (-1.342184662819, -1.402877211571, 1.547661781311, 0.745113432407)
    _e = -0.319332510233
    _scale, _offset = 977.555715042962, 579.184642915415
    _rate_div, _T0, _T_div = 10.819778284410, 20.0, 480.0

    def __init__(self, name=""):
        super().__init__(name)
        mn = [0.0, 0.0, self._T0, self._offset]
        mx = [1.0, self._rate_div, self._T0 + self._T_div, self._offset + self._scale]
        self._table = np.concatenate([[7, 4], -np.array(self._A).ravel(), -np.array(self._a),
                                      -np.array(self._C).ravel(), -np.array(self._c), self._E, [self._e, 1.0],
                                      mn, mx]).astype(np.float64)

    def setParameters(self, filename=None):
        pass


class Material:
    """Material attributes of Material.h:49-56."""

    def __init__(self, name=""):
        self.name = name
        self.youngModulus = 0.0
        self.poissonRatio = 0.0
        self.density = 0.0
        self.heatCapacity = 0.0
        self.taylorQuinney = 0.0
        self.T0 = 0.0
        self.dilatation = 0.0
        self.thermalConductivity = 0.0
        self._law = None

    def setHardeningLaw(self, law):
        self._law = law

    def getHardeningLaw(self):
        return self._law

    def getBulkModulus(self):
        return self.youngModulus / (3.0 * (1.0 - 2.0 * self.poissonRatio))

    def getShearModulus(self):
        return self.youngModulus / (2.0 * (1.0 + self.poissonRatio))

    def getTwoShearModulus(self):
        """Material.h:113."""
        return self.youngModulus / (1.0 + self.poissonRatio)

    def getLambda(self):
        """Material.h:141."""
        return self.poissonRatio * self.youngModulus / ((1.0 + self.poissonRatio) * (1.0 - 2.0 * self.poissonRatio))

    def getHeatFraction(self):
        """Material.h:168: Taylor-Quinney / (rho Cp)."""
        return self.taylorQuinney / (self.density * self.heatCapacity)

    def isYield(self):
        """Material::isYield (Material.C:170-175): the law's flag, false for
        ElasticLaw."""
        return self._law is not None and not isinstance(self._law, ElasticLaw)

    def setColor(self, *rgb):
        """Display colour (Material::setColor); not used by the solve."""
        self.color = tuple(float(c) for c in rgb)

    def descriptor(self):
        if self._law is None:
            raise DnlgError(_lib.E_INVALID, f"material {self.name!r} has no hardening law")
        d = _lib.MaterialDesc()
        d.young, d.poisson, d.density = self.youngModulus, self.poissonRatio, self.density
        d.heat_capacity, d.taylor_quinney, d.T0 = self.heatCapacity, self.taylorQuinney, self.T0
        d.law = self._law.law
        for k in range(8):
            d.p[k] = self._law.p[k]
        if isinstance(self._law, TabularLaw):
            f = self._law.getFunction()
            if f is None:
                raise DnlgError(_lib.E_INVALID, f"tabular law {self._law.name!r} has no function")
            d._table = np.ascontiguousarray(np.array(f._pts, dtype=np.float64).reshape(-1))
            d.table = _lib.as_dp(d._table)
            d.n_table = len(f._pts)
        if isinstance(self._law, ANN2SigLaw):
            if self._law._table is None:
                raise DnlgError(_lib.E_INVALID, f"network law {self._law.name!r} has no parameters")
            d._table = np.ascontiguousarray(self._law._table)
            d.table = _lib.as_dp(d._table)
            d.n_table = len(d._table)
        return d


class RampFunction:
    """RampFunction (RampFunction.h:31-39): setFunction(type, min, max)."""
    Unknown, Linear, Sinus, Constant, Erf, Soft = range(6)

    def __init__(self, name=""):
        self.name = name
        self.type = self.Unknown
        self.min = 0.0
        self.max = 1.0

    def setFunction(self, type_, mn, mx):
        self.type, self.min, self.max = int(type_), float(mn), float(mx)

    set = setFunction


class _Boundary:
    kind = None

    def __init__(self, name=""):
        self.name = name
        self.value = [0.0, 0.0, 0.0]
        self.function = None

    def setValue(self, x, y=None, z=None):
        if y is None:
            x, y, z = x(0), x(1), x(2)
        self.value = [float(x), float(y), float(z)]

    def getValue(self):
        return Vec3D(*self.value)

    def setFunction(self, function):
        self.function = function

    def descriptor(self):
        b = _lib.BoundaryDesc()
        b.kind = self.kind
        for k in range(3):
            b.value[k] = self.value[k]
        if self.function is not None and self.kind == _lib.BC_SPEED:
            b.ramp, b.ramp_min, b.ramp_max = self.function.type, self.function.min, self.function.max
        return b


class BoundaryRestrain(_Boundary):
    """BoundaryRestrain (BoundaryRestrain.h:49): setValue(rx, ry, rz)."""
    kind = _lib.BC_RESTRAIN


class BoundarySpeed(_Boundary):
    """BoundarySpeed (BoundarySpeed.h:49): setValue(vx, vy, vz), setFunction."""
    kind = _lib.BC_SPEED


class Explicit:
    """Explicit solver settings (Solver.h:73-78, Explicit.h:56).  Defaults are
    the reference's: Courant step every 10 increments, safety 0.9,
    rho_b = 0.8182 (Solver.h:30-33, configuration.dnl:9-11, Explicit.h:16).
    ``setPrecision("mixed")`` selects FP32 strain / stress / forces with the
    radial-return Newton solve, kinematics and accumulators in FP64;
    ``"mixed-rr32"`` also runs the radial return in FP32."""
    Courant, PowerIteration, PowerIterationUnder = 0, 1, 2

    def __init__(self, name=""):
        self.name = name
        self.startTime = 0.0
        self.endTime = 0.0
        self.timeStepMethod = self.Courant
        self.computeTimeStepFrequency = 10
        self.timeStepSafetyFactor = 0.9
        self.dissipation = 0.8182
        self.precision = _lib.PREC_FP64
        self.trackTensors = True
        self.nrMaxIterations = 0
        self.model = None

    def setTimes(self, start, end):
        self.startTime, self.endTime = float(start), float(end)

    def setTimeStepMethod(self, method):
        self.timeStepMethod = int(method)

    def setComputeTimeStepFrequency(self, frequency):
        self.computeTimeStepFrequency = int(frequency)

    def setTimeStepSafetyFactor(self, factor):
        self.timeStepSafetyFactor = float(factor)

    def setIncrements(self, start, stop):
        """Solver::setIncrements (Solver.C:175-180): kept for scripts; the
        explicit loop runs on time only (Explicit.C:184) and never consults
        the increment bounds, so neither does this engine."""
        self.startIncrement, self.endIncrement = int(start), int(stop)

    def setDissipation(self, rho_b):
        if rho_b < 0.0 or rho_b > 1.0:
            raise DnlgError(_lib.E_INVALID, "dissipation value out of range [0.0, 1.0]")
        self.dissipation = float(rho_b)

    def setPrecision(self, precision):
        p = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED,
             "mixed-rr32": _lib.PREC_MIXED_RR32}.get(str(precision).lower(), precision)
        if p not in (_lib.PREC_FP64, _lib.PREC_MIXED, _lib.PREC_MIXED_RR32):
            raise DnlgError(_lib.E_INVALID, f"unknown precision {precision!r}")
        self.precision = p

    def setTrackTensors(self, flag):
        self.trackTensors = bool(flag)

    def setMaxNewtonIterations(self, n):
        """Iterations of the radial return's Newton solve before the job
        stops with NR_DIVERGED (0: the reference's 250, Element.C:560-581).
        No counterpart in the reference, whose limit is fixed."""
        if int(n) < 0:
            raise DnlgError(_lib.E_INVALID, "the Newton iteration limit must be >= 0")
        self.nrMaxIterations = int(n)

    def getTimeStep(self):
        return self.model.getTimeStep() if self.model else 0.0

    def params(self):
        p = _lib.SolverParams()
        p.t_start, p.t_end = self.startTime, self.endTime
        p.rho_b, p.safety = self.dissipation, self.timeStepSafetyFactor
        p.dt_method, p.dt_freq = self.timeStepMethod, self.computeTimeStepFrequency
        p.stress_method = _lib.STRESS_NR
        p.precision = self.precision
        p.track_tensors = 1 if self.trackTensors else 0
        p.nr_max_iter = self.nrMaxIterations
        return p


class Context:
    """One engine context (device + stream).  ``Context.default()`` is shared
    by models that do not pass their own."""
    _default = {}

    def __init__(self, device=0):
        self.lib = _lib.load()
        self.handle = ctypes.c_void_p()
        _lib.check(self.lib.dnlg_ctx_create(int(device), ctypes.byref(self.handle)))
        self.device = device
        self._models = weakref.WeakSet()  # models whose engine job lives in this context
        # the measured schedule table next to the library (dnlPython.autotune)
        if os.environ.get("DNLG_SCHEDULES", "1") != "0":
            from . import autotune
            autotune.load(self)

    @classmethod
    def default(cls, device=0):
        if device not in cls._default:
            cls._default[device] = cls(device)
        return cls._default[device]

    def check(self, status):
        return _lib.check(status, self.handle)

    def set_timing(self, on):
        self.check(self.lib.dnlg_ctx_set_timing(self.handle, 1 if on else 0))

    def reset_timing(self):
        self.check(self.lib.dnlg_ctx_reset_timing(self.handle))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self.check(self.lib.dnlg_ctx_kernel_time(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def add_schedule(self, elem_type, precision, min_elems=0, max_elems=0, **fields):
        """A rule of the launch-schedule table (dnlg_ctx_add_schedule)."""
        s = _lib.Schedule(**{k: int(v) for k, v in fields.items()})
        self.check(self.lib.dnlg_ctx_add_schedule(self.handle, int(elem_type), int(precision), int(min_elems),
                                                  int(max_elems), ctypes.byref(s)))

    def clear_schedules(self):
        self.check(self.lib.dnlg_ctx_clear_schedules(self.handle))

    def close(self):
        """Destroys the context; the engine jobs of its models are released
        first (a job may not outlive its context)."""
        if self.handle:
            for m in list(self._models):
                m.close()
            self.lib.dnlg_ctx_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class _NodeView:
    """Node proxy for ``getNodeByNum`` (coords/disp/speed as Vec3D)."""

    def __init__(self, model, number):
        self._m, self.number = model, number

    def _row(self, what):
        arr = self._m._node_array(what)
        return Vec3D(*arr[:, self._m._node_index(self.number)])

    @property
    def coords(self):
        return self._row("coords")

    @property
    def disp(self):
        return self._row("disp")

    @property
    def speed(self):
        return self._row("speed")

    @property
    def mass(self):
        return float(self._m.getNodalMasses()[self._m._node_index(self.number)])


class _ElementView:
    """Element proxy for ``getElementByNum``: its number, node numbers in
    connectivity order, and integration-point values
    (Element::getIntPointValue, Element.C:1195-1238)."""

    def __init__(self, model, number):
        self._m, self.number = model, number
        self._index = model._element_index(number)

    @property
    def nodes(self):
        m = self._m
        return [int(m._node_sorted[k]) for k in m._arrays.conn[self._index]]

    def getNumberOfNodes(self):
        return NODES_PER_ELEMENT[self._m._elem_type]

    def getNumberOfIntegrationPoints(self):
        return IPS_PER_ELEMENT[self._m._elem_type]

    def getIntPointValue(self, field, ip):
        """A scalar integration-point field at IP `ip` of this element."""
        v = self._m.getIntPointField(field)
        if v.ndim != 2:
            raise DnlgError(_lib.E_INVALID, "field must be scalar")
        return float(v[self._index, ip])


class _Parallel:
    """Parallel (Parallel.h:45-80): core count is irrelevant on the GPU."""

    def setCores(self, n):
        self.cores = int(n)


class VtkInterface:
    """VtkInterface (VtkInterface.C): the point-data fields of the result
    files.  The default list is the VtkFields setting of
    Sources/dnlSettings/configuration.dnl, labels that name no field left out
    (VtkInterface::initFields)."""

    DEFAULT_LABELS = ("Stress", "Strain", "PlasticStrain", "vonMises", "yield", "pressure", "plasticStrain",
                      "plasticStrainRate", "gamma", "gammaCumulate", "temperature", "speed", "displacement",
                      "displacementIncrement")

    def __init__(self, name=""):
        self.name = name
        self._fields = [f for f in (Field.getField(n) for n in self.DEFAULT_LABELS) if f != -1]

    def existField(self, field):
        return self._fields.index(int(field)) if int(field) in self._fields else -1

    def addField(self, field):
        if self.existField(field) == -1:
            self._fields.append(int(field))

    def removeField(self, field):
        k = self.existField(field)
        if k != -1:
            del self._fields[k]

    def getNumberOfFields(self):
        return len(self._fields)


class SvgInterface:
    """SvgInterface (SvgInterface.h): SVG contour plots of a field.  Not
    provided: the view and legend settings are accepted so that a sample
    script runs its solve, and write() warns that no SVG file was written."""

    def __init__(self, name=""):
        self.name = name

    def __getattr__(self, attr):
        if attr.startswith(("set", "rotate", "reset")):
            return lambda *args, **kwargs: None
        raise AttributeError(attr)

    def write(self, fileName, field=None):
        warnings.warn(f"SvgInterface: SVG output is not provided, {fileName} was not written", RuntimeWarning,
                      stacklevel=2)


class HistoryFile:
    """HistoryFile (HistoryFile.h, HistoryFile.C): time series of node,
    element and global values, one row per save time, written by the engine
    while it solves (rows taken on the GPU after each cycle, written to the
    file between launch sequences).  Items are bound when added, like the
    reference's (the set's members at that moment)."""

    def __init__(self, name=""):
        self.name = name
        self._file_name = ""
        self._start, self._stop, self._save = 0.0, 0.0, 0.0
        self._items = []  # (kind, field, node/element number, intPt, header name)
        self._model = None
        self._id = None

    def setFileName(self, fileName):
        self._file_name = str(fileName)

    def getFileName(self):
        return self._file_name

    def setSaveTime(self, *args):
        """setSaveTime(saveTime) or setSaveTime(startTime, stopTime, saveTime)
        (HistoryFile.C:131-148)."""
        if len(args) == 1:
            self._start, self._stop, self._save = 0.0, -1.0, float(args[0])
        elif len(args) == 3:
            self._start, self._stop, self._save = float(args[0]), float(args[1]), float(args[2])
        else:
            raise TypeError("setSaveTime(saveTime) or setSaveTime(startTime, stopTime, saveTime)")

    def getSaveTime(self):
        return self._save

    def getStartTime(self):
        return self._start

    def getStopTime(self):
        return self._stop

    def add(self, *args):
        """add(nodeSet, field), add(elementSet, intPt, field) or add(field)
        (HistoryFile.C:186-280): one item per set member, named
        label_n<number> / label[_intPt]_e<number> when the set has more than
        one member."""
        if self._model is not None and self._model._job is not None:
            raise DnlgError(_lib.E_STATE, "items must be added before the history file is attached to a solved model")
        if len(args) == 1:
            f = int(args[0])
            self._items.append((_lib.H_GLOBAL, f, 0, 0, Field.getVtklabel(f)))
        elif len(args) == 2:
            nodes, f = args[0].numbers(), int(args[1])
            base = Field.getVtklabel(f)
            for n in nodes:
                self._items.append((_lib.H_NODE, f, int(n), 0, f"{base}_n{n}" if len(nodes) > 1 else base))
        elif len(args) == 3:
            elems, ip, f = list(args[0]._numbers), int(args[1]), int(args[2])
            base = Field.getVtklabel(f) + (f"_{ip}" if ip > 0 else "")
            for e in elems:
                self._items.append((_lib.H_ELEMENT, f, int(e), ip, f"{base}_e{e}" if len(elems) > 1 else base))
        else:
            raise TypeError("add(nodeSet, field), add(elementSet, intPt, field) or add(field)")

    def _attach(self, model):
        """Registers the file with the model's engine job (before its first
        solve)."""
        if not self._file_name:
            raise DnlgError(_lib.E_INVALID, f"No file name specified for object : {self.name}")
        n = len(self._items)
        items = (_lib.HistoryItem * max(1, n))()
        names = (ctypes.c_char_p * max(1, n))()
        for k, (kind, f, num, ip, label) in enumerate(self._items):
            if kind == _lib.H_NODE:
                idx = int(model._index_of(model._node_sorted, np.array([num]), "history node")[0])
            elif kind == _lib.H_ELEMENT:
                idx = int(model._index_of(model._elem_sorted, np.array([num]), "history element")[0])
            else:
                idx = 0
            items[k] = _lib.HistoryItem(kind, f, idx, ip, 0)
            names[k] = label.encode()
        fid = ctypes.c_int32()
        ctx = model.context
        ctx.check(ctx.lib.dnlg_job_add_history(model._job, self._file_name.encode(), self._start, self._save, items,
                                               n, names, ctypes.byref(fid)))
        self._id = fid.value

    def getRows(self):
        """The rows written so far, [rows, 1 + items] (time first)."""
        if self._model is None or self._id is None:
            return np.zeros((0, 1 + len(self._items)))
        m = self._model
        lib = m.context.lib
        n = ctypes.c_int64()
        m.context.check(lib.dnlg_job_history_rows(m._job, self._id, None, 0, ctypes.byref(n)))
        out = np.zeros((n.value, 1 + len(self._items)))
        if n.value:
            m.context.check(lib.dnlg_job_history_rows(m._job, self._id, _lib.as_dp(out), n.value, ctypes.byref(n)))
        return out

    def getItemNames(self):
        return [it[4] for it in self._items]


class DynELA:
    """Model facade (DynELA.h:40-111)."""

    def __init__(self, name="", context=None):
        self.name = name
        self._ctx = context
        self._node_num = []
        self._node_xyz = []
        self._elem_num = []
        self._elem_conn = []
        self._default_element = None
        self._elem_type = None
        self._solve_seconds = 0.0  # wall time inside solve calls (displayEstimatedEnd)
        self._last_display = (0.0, 0.0)
        self._material = None
        self._solver = None
        self._initial = []   # (boundary, node numbers) in attach order
        self._constant = []
        self._job = None
        self._arrays = None
        self.parallel = _Parallel()
        self.saveTimes = None
        self._stress_method = "NewtonRaphson"
        self._histories = []
        self.dataFile = VtkInterface()
        self._results_started = False

    # ---- nodes and elements -------------------------------------------------
    def createNode(self, number, x, y=None, z=None):
        if y is None:
            x, y, z = x(0), x(1), x(2)
        self._node_num.append(np.array([int(number)], dtype=np.int64))
        self._node_xyz.append(np.array([[float(x), float(y), float(z)]], dtype=np.float64))
        self._invalidate()
        return True

    def createNodes(self, numbers, xyz):
        """Bulk node ingestion: numbers [N], xyz [N, 3] (or [N, 2])."""
        xyz = np.asarray(xyz, dtype=np.float64)
        if xyz.shape[1] == 2:
            xyz = np.concatenate([xyz, np.zeros((xyz.shape[0], 1))], axis=1)
        self._node_num.append(np.asarray(numbers, dtype=np.int64).ravel())
        self._node_xyz.append(xyz)
        self._invalidate()
        return True

    def setDefaultElement(self, type_):
        self._default_element = int(type_)

    def createElement(self, number, *nodes):
        if self._default_element is None:
            raise DnlgError(_lib.E_INVALID, "setDefaultElement must be called first")
        self._check_type(self._default_element)
        if len(nodes) == 1 and hasattr(nodes[0], "__len__"):
            nodes = tuple(nodes[0])
        if len(nodes) != NODES_PER_ELEMENT[self._default_element]:
            raise DnlgError(_lib.E_INVALID, f"element {number} needs "
                            f"{NODES_PER_ELEMENT[self._default_element]} nodes")
        self._elem_num.append(np.array([int(number)], dtype=np.int64))
        self._elem_conn.append(np.array([nodes], dtype=np.int64))
        self._invalidate()
        return True

    def createElements(self, numbers, conn):
        """Bulk element ingestion with node numbers: conn [NE, nodes]."""
        if self._default_element is None:
            raise DnlgError(_lib.E_INVALID, "setDefaultElement must be called first")
        self._check_type(self._default_element)
        conn = np.asarray(conn, dtype=np.int64)
        if conn.ndim != 2 or conn.shape[1] != NODES_PER_ELEMENT[self._default_element]:
            raise DnlgError(_lib.E_INVALID, "connectivity has the wrong shape")
        self._elem_num.append(np.asarray(numbers, dtype=np.int64).ravel())
        self._elem_conn.append(conn)
        self._invalidate()
        return True

    def _check_type(self, t):
        if t not in NODES_PER_ELEMENT:
            raise DnlgError(_lib.E_INVALID, f"element type {t} is not supported by the engine")
        if self._elem_type is not None and self._elem_type != t:
            raise DnlgError(_lib.E_INVALID, "a model holds a single element family (Model.C:328-357)")
        self._elem_type = t

    def getNodesNumber(self):
        return int(sum(len(a) for a in self._node_num))

    def getElementsNumber(self):
        return int(sum(len(a) for a in self._elem_num))

    # ---- sets, materials, BCs, solver ---------------------------------------
    def add(self, obj, *args):
        if isinstance(obj, NodeSet):
            obj._numbers.extend(self._range(*args))
        elif isinstance(obj, ElementSet):
            obj._numbers.extend(self._range(*args))
        elif isinstance(obj, Material):
            self._material = obj
        elif isinstance(obj, Explicit):
            self._solver = obj
            obj.model = self
        elif isinstance(obj, HistoryFile):
            # Model::add(HistoryFile *) (Model.C:927-939)
            if obj._model is not None and obj._model is not self:
                raise DnlgError(_lib.E_INVALID, "the history file belongs to another model")
            obj._model = self
            self._histories.append(obj)
            if self._job is not None:
                obj._attach(self)
            return
        else:
            raise DnlgError(_lib.E_INVALID, f"cannot add a {type(obj).__name__} to the model")
        self._invalidate()

    @staticmethod
    def _range(start=-1, end=-1, increment=1):
        """Model::add(set, start, end, increment) (Model.C:181-233)."""
        if start == -1:
            return []
        if end == -1:
            return [int(start)]
        return list(range(int(start), int(end) + 1, int(increment)))

    def addMaterial(self, material):
        self._material = material

    def attachInitialBC(self, boundary, nodeSet):
        self._initial.append((boundary, nodeSet.numbers()))
        self._invalidate()

    def attachConstantBC(self, boundary, nodeSet):
        self._constant.append((boundary, nodeSet.numbers()))
        self._invalidate()

    def setSaveTimes(self, start, end, increment):
        self.saveTimes = (float(start), float(end), float(increment))

    def setStressComputation(self, method):
        """DynELA::setStressComputation (DynELA.C:1406-1430): "NewtonRaphson"
        (default) or "Direct" (one linearised return, in FP64 on the FP64 and
        mixed paths; not "mixed-rr32")."""
        if method not in ("NewtonRaphson", "Direct"):
            raise DnlgError(_lib.E_INVALID, f"unknown stress computation method {method!r}")
        self._invalidate()
        self._stress_method = method

    # ---- assembly of the engine descriptors -----------------------------------
    def _invalidate(self):
        if self._job is not None:
            raise DnlgError(_lib.E_STATE, "the model was modified after initSolve/solve")
        self._arrays = None

    def _node_tables(self):
        nums = np.concatenate(self._node_num) if self._node_num else np.zeros(0, np.int64)
        xyz = np.concatenate(self._node_xyz) if self._node_xyz else np.zeros((0, 3))
        order = np.argsort(nums, kind="stable")
        nums, xyz = nums[order], xyz[order]
        if len(nums) > 1 and np.any(nums[1:] == nums[:-1]):
            raise DnlgError(_lib.E_INVALID, "duplicate node number")
        return nums, xyz

    def _index_of(self, sorted_nums, numbers, what):
        idx = np.searchsorted(sorted_nums, numbers)
        idx = np.clip(idx, 0, max(len(sorted_nums) - 1, 0))
        if len(numbers) and (len(sorted_nums) == 0 or np.any(sorted_nums[idx] != numbers)):
            raise DnlgError(_lib.E_INVALID, f"{what} does not exist in the model")
        return idx

    def descriptors(self):
        """Flat engine descriptors of the model (ModelArrays)."""
        if self._arrays is not None:
            return self._arrays
        if self._elem_type is None or not self._elem_num:
            raise DnlgError(_lib.E_INVALID, "model has no elements")
        if self._material is None:
            raise DnlgError(_lib.E_INVALID, "no material attached to the elements")
        if self._solver is None:
            raise DnlgError(_lib.E_INVALID, "no solver added to the model")
        nums, xyz = self._node_tables()
        enums = np.concatenate(self._elem_num)
        econn = np.concatenate(self._elem_conn)
        order = np.argsort(enums, kind="stable")
        enums, econn = enums[order], econn[order]
        conn = self._index_of(nums, econn.ravel(), "element node").reshape(econn.shape).astype(np.int32)
        n = len(nums)
        bnds, lists = [], []
        for which in (self._initial, self._constant):
            rows, keys, vals = [], [], []
            for k, (b, members) in enumerate(which):
                bi = len(bnds)
                bnds.append(b.descriptor())
                idx = self._index_of(nums, members, "boundary node")
                rows.append(idx)
                keys.append(np.full(len(idx), k, dtype=np.int64))
                vals.append(np.full(len(idx), bi, dtype=np.int32))
            if rows:
                r, kk, v = np.concatenate(rows), np.concatenate(keys), np.concatenate(vals)
                o = np.lexsort((kk, r))
                ptr = np.zeros(n + 1, dtype=np.int64)
                np.add.at(ptr, r + 1, 1)
                lists.append((np.cumsum(ptr), v[o]))
            else:
                lists.append(None)
        params = self._solver.params()
        params.stress_method = _lib.STRESS_DIRECT if self._stress_method == "Direct" else _lib.STRESS_NR
        self._arrays = _lib.ModelArrays(self._elem_type, xyz.T.copy(), conn, self._material.descriptor(), bnds,
                                        lists[0], lists[1], params)
        self._node_sorted = nums
        self._elem_sorted = enums
        return self._arrays

    # ---- engine -------------------------------------------------------------
    @property
    def context(self):
        if self._ctx is None:
            self._ctx = Context.default()
        return self._ctx

    def initSolve(self):
        """Model::initSolve (Model.C:360-521): creates the engine job."""
        if self._job is not None:
            return True
        a = self.descriptors()
        ctx = self.context
        job = ctypes.c_void_p()
        ctx.check(ctx.lib.dnlg_job_create(ctx.handle, ctypes.byref(a.mesh), ctypes.byref(a.material),
                                          ctypes.byref(a.bcs), ctypes.byref(a.params), ctypes.byref(job)))
        self._job = job
        ctx._models.add(self)
        for h in self._histories:
            h._attach(self)
        if self.saveTimes is not None:
            # DynELA::setSaveTimes, result files <name><index>.vtk (DynELA.C:59, 688-710)
            start, end, inc = self.saveTimes
            fields = np.array(self.dataFile._fields, dtype=np.int32)
            ctx.check(ctx.lib.dnlg_job_set_result_files(job, self.name.encode(), start, end, inc,
                                                        _lib.as_i32p(fields), len(fields)))
        return True

    def solve(self, upTo=None):
        """DynELA::solve (DynELA.C:727-772) / Model::solve(upTo).  With save
        times set, the result files of the reference: the initial
        configuration before the first solve (writeResultFile), one after
        every cycle that reaches the next save time, and the final state
        once the end time is reached (writeVTKFile)."""
        self.initSolve()
        t = self._solver.endTime if upTo is None or upTo < 0 else float(upTo)
        ctx = self.context
        if self.saveTimes is not None and not self._results_started:
            self._results_started = True
            self.writeResultFile()
        t0 = time.perf_counter()
        try:
            ctx.check(ctx.lib.dnlg_job_solve(self._job, ctypes.c_double(t)))
        finally:
            self._solve_seconds += time.perf_counter() - t0
        if self.saveTimes is not None and self.getCurrentTime() >= self._solver.endTime:
            self.writeVTKFile()
        return True

    def writeResultFile(self):
        """DynELA::writeResultFile (DynELA.C:489-503)."""
        self.initSolve()
        self.context.check(self.context.lib.dnlg_job_write_result_file(self._job, 0))

    def writeVTKFile(self):
        """DynELA::writeVTKFile (DynELA.C:688-710)."""
        self.initSolve()
        self.context.check(self.context.lib.dnlg_job_write_result_file(self._job, 1))

    def close(self):
        if self._job is not None:
            if self.context.handle:
                self.context.lib.dnlg_job_destroy(self._job)
            self._job = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- results --------------------------------------------------------------
    def _require_job(self):
        if self._job is None:
            raise DnlgError(_lib.E_STATE, "call initSolve() or solve() first")

    def _node_array(self, what):
        self._require_job()
        n = self.getNodesNumber()
        out = np.zeros((3, n))
        ptrs = {"coords": None, "disp": None, "speed": None}
        ptrs[what] = _lib.as_dp(out)
        self.context.check(self.context.lib.dnlg_job_read_nodes(self._job, ptrs["coords"], ptrs["disp"],
                                                                ptrs["speed"]))
        return out

    def _node_index(self, number):
        if self._arrays is None:
            self.descriptors()
        return int(self._index_of(self._node_sorted, np.array([number]), "node")[0])

    def getNodalCoordinates(self):
        """[3, N] current coordinates, nodes in increasing number order."""
        return self._node_array("coords")

    def getNodalDisplacements(self):
        return self._node_array("disp")

    def getNodalSpeeds(self):
        return self._node_array("speed")

    def getNodalMasses(self):
        self._require_job()
        out = np.zeros(self.getNodesNumber())
        self.context.check(self.context.lib.dnlg_job_read_mass(self._job, _lib.as_dp(out)))
        return out

    def getIntPointField(self, field):
        """[n_elems, n_ip] (scalars) or [6, n_elems, n_ip] (tensors)."""
        self._require_job()
        key = getattr(Field, field, None) if isinstance(field, str) else field
        if key not in _IP_FIELDS:
            raise DnlgError(_lib.E_INVALID, f"no integration-point field {field!r}")
        f = _IP_FIELDS[key]
        ne = self.getElementsNumber()
        nip = IPS_PER_ELEMENT[self._elem_type]
        ncomp = 6 if f in _lib.TENSOR_FIELDS else 1
        out = np.zeros(ncomp * ne * nip)
        self.context.check(self.context.lib.dnlg_job_read_ip(self._job, f, _lib.as_dp(out)))
        return out.reshape((ncomp, ne, nip) if ncomp == 6 else (ne, nip))

    def getScalars(self):
        self._require_job()
        t, dt, inc = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self.context.check(self.context.lib.dnlg_job_scalars(self._job, ctypes.byref(t), ctypes.byref(dt),
                                                             ctypes.byref(inc)))
        return t.value, dt.value, inc.value

    def getCurrentTime(self):
        return self.getScalars()[0]

    def getTimeStep(self):
        return self.getScalars()[1]

    def getIncrement(self):
        return self.getScalars()[2]

    def getEnergies(self):
        """[kinetic, internal, dissipated plastic work, kinetic + internal]."""
        self._require_job()
        out = np.zeros(4)
        self.context.check(self.context.lib.dnlg_job_energies(self._job, _lib.as_dp(out)))
        return out

    def getTotalKineticEnergy(self):
        return float(self.getEnergies()[0])

    # ---- launch schedule (no counterpart in the reference, whose element
    # loops take fixed OpenMP chunks, Parallel.C:78-113) ---------------------
    def getSchedule(self):
        """The launch schedule the job runs with (dict: elem_block, waves,
        corr_batch, graph_cycles)."""
        self.initSolve()
        s = _lib.Schedule()
        self.context.check(self.context.lib.dnlg_job_get_schedule(self._job, ctypes.byref(s)))
        return s.as_dict()

    def setSchedule(self, **fields):
        """Set this job's schedule (fields not given keep the defaults)."""
        self.initSolve()
        s = _lib.Schedule(**{k: int(v) for k, v in fields.items()})
        self.context.check(self.context.lib.dnlg_job_set_schedule(self._job, ctypes.byref(s)))

    def autotune(self, cycles_per_trial=200, rounds=3):
        """Measure the candidate schedules on this job by advancing it
        (dnlg_job_autotune); returns (schedule dict, ms per cycle of the best
        schedule, ms per cycle of the starting one) and leaves the best
        schedule set."""
        self.initSolve()
        s, ms = _lib.Schedule(), (ctypes.c_double * 2)()
        self.context.check(self.context.lib.dnlg_job_autotune(self._job, int(cycles_per_trial), int(rounds),
                                                              ctypes.byref(s), ms))
        return s.as_dict(), ms[0], ms[1]

    # ---- checkpoint / restart (no counterpart in the reference, whose
    # Model::readData / writeData are commented out, Model.C:1031-1100) -------
    def getState(self):
        """The solver state as a dict of numpy arrays (dnlg_job_read_state):
        time, timeStep, increment, coords / disp_inc / speed / accel [3, N]
        (field0), stress [6, G], ip_scalars [9, G] (rows named by
        _lib.STATE_IP_NAMES), and strain / plastic_strain_tensor with tracked
        tensors."""
        self._require_job()
        n, g = self.getNodesNumber(), self.getElementsNumber() * IPS_PER_ELEMENT[self._elem_type]
        arr, rec = _lib.state_alloc(n, g, self._solver.trackTensors)
        self.context.check(self.context.lib.dnlg_job_read_state(self._job, ctypes.byref(rec)))
        arr.update(time=rec.time, timeStep=rec.time_step, increment=rec.increment)
        return arr

    def setState(self, state):
        """Continue this model from a state of getState (of this model, a
        copy of it built the same way, or the CPU oracle's): the next solve
        goes on from state['time'] (dnlg_job_write_state)."""
        self.initSolve()
        arr = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in state.items()
               if isinstance(v, np.ndarray)}
        rec = _lib.state_record(arr, float(state["time"]), float(state["timeStep"]), int(state["increment"]))
        self.context.check(self.context.lib.dnlg_job_write_state(self._job, ctypes.byref(rec)))

    def getTotalMass(self):
        return float(self.getNodalMasses().sum())

    def getNodeByNum(self, number):
        return _NodeView(self, number)

    def _element_index(self, number):
        if self._arrays is None:
            self.descriptors()
        return int(self._index_of(self._elem_sorted, np.array([number]), "element")[0])

    def getElementByNum(self, number):
        return _ElementView(self, number)

    def displayEstimatedEnd(self):
        """DynELA::displayEstimatedEnd (DynELA.C:506-538): the solve calls'
        wall time so far and the remaining time estimated from the rate
        since the previous display, in the reference's format."""
        t = self.getCurrentTime() if self._job is not None else 0.0
        if t <= 0.0:
            return
        elapsed = self._solve_seconds
        last_wall, last_t = self._last_display
        remaining = int((elapsed - last_wall) / (t - last_t) * (self._solver.endTime - t)) if t > last_t else 0

        def hms(sec):
            sec = int(sec)
            return f"{sec // 3600:02d}:{sec % 3600 // 60:02d}:{sec % 60:02d}"

        print(f"Current computational time {hms(elapsed)}")
        print(f"Estimated end of computation in {hms(remaining)}")
        self._last_display = (elapsed, t)

    def getNodalField(self, field):
        """Node::fieldScalar (Node.C:311-370) of every node in node-number
        order: nodal values, or IP values extrapolated to the nodes and
        averaged over each node's elements."""
        self._require_job()
        out = np.zeros(self.getNodesNumber())
        self.context.check(self.context.lib.dnlg_job_read_node_field(self._job, int(field), _lib.as_dp(out)))
        return out

    def getNodalValuesRange(self, field):
        """DynELA::getNodalValuesRange (DynELA.C:775-796): (min, max) of a
        scalar field over the nodes."""
        if int(field) in (Field.coords, Field.disp, Field.dispInc, Field.speed, Field.speedInc, Field.Strain,
                          Field.StrainInc, Field.PlasticStrain, Field.PlasticStrainInc, Field.Stress):
            raise DnlgError(_lib.E_INVALID, "field must be scalar")
        v = self.getNodalField(field)
        return float(v.min()), float(v.max())

    def getGlobalBox(self):
        """DynELA::getGlobalBox (DynELA.C:663-685): (min, max) corners of the
        nodes' bounding box, current coordinates once solved."""
        c = self.getNodalCoordinates() if self._job is not None else self._node_tables()[1].T
        return Vec3D(*c.min(axis=1)), Vec3D(*c.max(axis=1))

    # ---- geometry transforms of the mesh (before the first solve) ---------------
    def _transform(self, fn, nodeSet):
        nums, xyz = self._node_tables()
        sel = np.ones(len(nums), bool) if nodeSet is None else np.isin(nums, nodeSet.numbers())
        xyz = xyz.copy()
        xyz[sel] = fn(xyz[sel])
        self._invalidate()
        self._node_num, self._node_xyz = [nums], [xyz]

    def translate(self, vector, nodeSet=None):
        """DynELA::translate (DynELA.C:542-552)."""
        v = np.array([vector(0), vector(1), vector(2)])
        self._transform(lambda x: x + v, nodeSet)

    def scale(self, value, nodeSet=None):
        """DynELA::scale (DynELA.C:554-578): a number or a Vec3D per axis."""
        v = np.array([value(0), value(1), value(2)]) if isinstance(value, Vec3D) else float(value)
        self._transform(lambda x: x * v, nodeSet)

    def rotate(self, axis, angle, nodeSet=None):
        """DynELA::rotate (DynELA.C:580-661), angle in degrees: about a Vec3D
        axis (right-handed, Rodrigues' formula), or about "X" / "Y" / "Z";
        the reference's "X" and "Y" matrices turn by -angle, its "Z" matrix
        by +angle, and this keeps that."""
        if isinstance(axis, str):
            a = axis.upper()
            if a not in ("X", "Y", "Z"):
                return
            k = np.eye(3)["XYZ".index(a)]
            th = np.radians(angle) * (1.0 if a == "Z" else -1.0)
        else:
            k = np.array([axis(0), axis(1), axis(2)], dtype=np.float64)
            k = k / np.linalg.norm(k)
            th = np.radians(angle)
        K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
        R = np.cos(th) * np.eye(3) + np.sin(th) * K + (1.0 - np.cos(th)) * np.outer(k, k)
        self._transform(lambda x: x @ R.T, nodeSet)


def batch_solve(models, upTo=None, context=None, raise_on_error=True, max_cycles=None):
    """Advance several models together (one launch per phase and cycle).

    A job that fails (negative Jacobian, Newton divergence, non-finite state)
    stops; the others run to their target time.  With raise_on_error the
    first failure raises after the batch; otherwise the per-model status
    codes are returned (0 = OK).  max_cycles: return after about that many
    cycles even if some model has not reached its target (check
    getCurrentTime)."""
    ctx = context or models[0].context
    for m in models:
        if m.context is not ctx:
            raise DnlgError(_lib.E_INVALID, "all models of a batch must share one context")
        m.initSolve()
    n = len(models)
    handles = (ctypes.c_void_p * n)(*[m._job.value for m in models])
    t = np.array([m._solver.endTime if upTo is None else float(upTo) for m in models], dtype=np.float64)
    if max_cycles is None:
        st = ctx.lib.dnlg_batch_solve(ctx.handle, handles, n, _lib.as_dp(t))
    else:
        st = ctx.lib.dnlg_batch_advance(ctx.handle, handles, n, _lib.as_dp(t), int(max_cycles))
    statuses = [ctx.lib.dnlg_job_status(m._job) for m in models]
    if raise_on_error:
        ctx.check(st)
    elif st != _lib.OK and all(s == _lib.OK for s in statuses):
        ctx.check(st)  # a failure of the batch itself, not of a job
    return statuses


class Server:
    """Continuous-batching serving loop over one context (dnlg_server_*).

    ``submit(model)`` queues a model built with this API and returns a ticket;
    ``step(cycles)`` admits queued jobs whose device footprint fits in the
    memory budget, advances every running job by about ``cycles`` cycles and
    retires finished ones; ``take(ticket)`` returns the finished model (its
    getters read the engine job as after ``solve()``) together with its
    status.  Results of a job do not depend on the jobs it ran beside.
    cycle_budget > 0 stops a job still short of its end time after that many
    increments (status E_CAPACITY)."""

    QUEUED, RUNNING, DONE = 0, 1, 2

    def __init__(self, context=None, hbm_budget=0, max_jobs=0, cycle_budget=0):
        self.ctx = context or Context.default()
        self.handle = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.dnlg_server_create(self.ctx.handle, int(hbm_budget), int(max_jobs),
                                                       ctypes.byref(self.handle)))
        if cycle_budget:
            self.ctx.check(self.ctx.lib.dnlg_server_set_cycle_budget(self.handle, int(cycle_budget)))
        self._models = {}

    def submit(self, model):
        if model._job is not None:
            raise DnlgError(_lib.E_STATE, "the model already has an engine job")
        a = model.descriptors()
        t = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.dnlg_server_submit(self.handle, ctypes.byref(a.mesh), ctypes.byref(a.material),
                                                       ctypes.byref(a.bcs), ctypes.byref(a.params),
                                                       ctypes.byref(t)))
        self._models[t.value] = model
        return t.value

    def step(self, cycles=64):
        """Returns (running, queued) after the step."""
        r, q = ctypes.c_int32(), ctypes.c_int32()
        self.ctx.check(self.ctx.lib.dnlg_server_step(self.handle, int(cycles), ctypes.byref(r), ctypes.byref(q)))
        return r.value, q.value

    def poll(self, ticket):
        """(state, status): state QUEUED / RUNNING / DONE."""
        st, code = ctypes.c_int32(), ctypes.c_int32()
        self.ctx.check(self.ctx.lib.dnlg_server_poll(self.handle, int(ticket), ctypes.byref(st), ctypes.byref(code)))
        return st.value, code.value

    def take(self, ticket):
        """(model, status) of a finished job; the model owns the job from now on."""
        job = ctypes.c_void_p()
        status = self.ctx.lib.dnlg_server_take(self.handle, int(ticket), ctypes.byref(job))
        model = self._models.pop(int(ticket))
        if job.value:
            model._job = job
            model._ctx = self.ctx
            self.ctx._models.add(model)
        elif status != _lib.OK:
            self.ctx.check(status)
        return model, status

    def close(self):
        if self.handle:
            self.ctx.lib.dnlg_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()
            self._models.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
