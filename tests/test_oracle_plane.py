"""Plane-strain 4-node quad (ElQua4N2D) in the CPU oracle (no GPU): the
element stiffness against closed forms, the lumped mass against rho x
area, free flight, and the elastic simple shear against the Jaumann closed
form -- the same pins as the 3D element's, for the plane element's
kinematics (F_zz = 1, no hoop terms, unit thickness)."""
import numpy as np
import pytest

import dnlPython as dnl
from oracle.bridge import F_STRESS, OracleJob, element_stiffness
from test_oracle_solve import elastic_law

PLANE = 1
E, NU = 200000.0, 0.3
FCT = E / ((1 + NU) * (1 - 2 * NU))


def _quad_model(xyz, conn, law, speed_sets=(), restrain_sets=(), initial=None, t_end=1e-5, safety=0.9):
    m = dnl.DynELA("plane")
    m.createNodes(np.arange(1, len(xyz) + 1), np.asarray(xyz, float))
    m.setDefaultElement(dnl.Element.ElQua4N2D)
    m.createElements(np.arange(1, len(conn) + 1), np.asarray(conn) + 1)
    es = dnl.ElementSet("all")
    m.add(es, 1, len(conn))
    mat = dnl.Material("m")
    mat.youngModulus, mat.poissonRatio, mat.density = 206e9, 0.3, 7830.0
    mat.heatCapacity, mat.taylorQuinney, mat.T0 = 460.0, 0.9, 20.0
    mat.setHardeningLaw(law)
    m.add(mat, es)
    for kind, sets in ((dnl.BoundaryRestrain, restrain_sets), (dnl.BoundarySpeed, speed_sets)):
        for vals, nodes in sets:
            b = kind("b")
            b.setValue(*vals)
            ns = dnl.NodeSet("n")
            ns._numbers.extend(int(n) + 1 for n in nodes)
            m.attachConstantBC(b, ns)
    if initial is not None:
        b = dnl.BoundarySpeed("v0")
        b.setValue(*initial)
        ns = dnl.NodeSet("all")
        ns._numbers.extend(range(1, len(xyz) + 1))
        m.attachInitialBC(b, ns)
    s = dnl.Explicit("s")
    s.setTimes(0.0, t_end)
    s.setTimeStepSafetyFactor(safety)
    m.add(s)
    return m


def test_plane_stiffness_closed_forms():
    """Rectangle a x b: symmetric, three rigid modes (two translations, one
    rotation); u = eps x gives sigma_xx = f (1 - nu) eps, sigma_yy = f nu eps
    (plane strain) and nodal forces sigma x edge / 2; simple shear u_x = g y
    gives sigma_xy = E / (1 + nu) g with the reference's (1 - 2 nu) shear
    diagonal (Material.C:104-111)."""
    a, b = 2.0, 3.0
    X = np.array([[0, 0, 0], [a, 0, 0], [a, b, 0], [0, b, 0]], float)
    K = element_stiffness(PLANE, X, E, NU)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() > -1e-9 * ev.max() and np.sum(ev < 1e-9 * ev.max()) == 3
    rot = np.stack([-X[:, 1], X[:, 0]], axis=1).ravel()
    assert np.abs(K @ rot).max() <= 1e-9 * np.abs(K).max() * b
    eps = 1e-3
    u = np.zeros(8)
    u[0::2] = eps * X[:, 0]
    f = (K @ u).reshape(4, 2)
    np.testing.assert_allclose(f[:, 0], np.where(X[:, 0] > 0, 1, -1) * FCT * (1 - NU) * eps * b / 2, rtol=1e-12)
    np.testing.assert_allclose(f[:, 1], np.where(X[:, 1] > 0, 1, -1) * FCT * NU * eps * a / 2, rtol=1e-12)
    g = 2e-3
    u = np.zeros(8)
    u[0::2] = g * X[:, 1]
    f = (K @ u).reshape(4, 2)
    sxy = E / (1 + NU) * g
    np.testing.assert_allclose(f[:, 0], np.where(X[:, 1] > 0, 1, -1) * sxy * a / 2, rtol=1e-12)
    np.testing.assert_allclose(f[:, 1], np.where(X[:, 0] > 0, 1, -1) * sxy * b / 2, rtol=1e-12)


def test_plane_mass_and_free_flight():
    """Lumped masses sum to rho x area (unit thickness, distorted mesh); a free
    body with speed v moves rigidly: x = x0 + v t, no stress."""
    xyz = np.array([[0, 0, 0], [1.0, 0.1, 0], [2.1, 0, 0], [0.1, 1.0, 0], [1.2, 1.1, 0], [2.0, 0.9, 0]]) * 0.01
    conn = [[0, 1, 4, 3], [1, 2, 5, 4]]
    v = (3.0, -2.0)
    m = _quad_model(xyz, conn, elastic_law(), initial=(v[0], v[1], 0.0), t_end=2e-5, safety=0.5)
    o = OracleJob(m.descriptors())
    assert o.solve(2e-5) == 0

    def area(q):
        x, y = xyz[q, 0], xyz[q, 1]
        return 0.5 * abs(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))

    assert o.mass().sum() == pytest.approx(7830.0 * sum(area(q) for q in conn), rel=1e-12)
    t, _, _ = o.scalars()
    _, d, sp = o.nodes()
    np.testing.assert_allclose(d[:2], np.array(v)[:, None] * t * np.ones((2, 6)), rtol=1e-12)
    np.testing.assert_allclose(sp[:2], np.array(v)[:, None] * np.ones((2, 6)), rtol=1e-10)
    assert np.abs(o.ip(F_STRESS)).max() < 1e-2


def test_plane_simple_shear_jaumann_closed_form():
    """One plane-strain quad in homogeneous simple shear (bottom fixed, top
    moving at V along x): the Jaumann solution s12 = G sin g,
    s11 - s22 = 2 G (1 - cos g), s33 = 0 (no in-plane coupling to zz)."""
    h = 1.0
    xyz = np.array([[0, 0, 0], [h, 0, 0], [h, h, 0], [0, h, 0]], float)
    E_, nu, rho = 206e9, 0.3, 7830.0
    c = np.sqrt(E_ * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))
    dt = 0.9 * 1.981818700333 * (h / np.sqrt(2) / c) / 2
    V = 2e-4 * h / dt
    t_end = 1.0 * h / V
    m = _quad_model(xyz, [[0, 1, 2, 3]], elastic_law(), restrain_sets=[((1, 1, 0), [0, 1]), ((0, 1, 0), [2, 3])],
                    speed_sets=[((V, 0.0, 0.0), [2, 3])], t_end=t_end)
    o = OracleJob(m.descriptors())
    assert o.solve(t_end) == 0
    cx, _, _ = o.nodes()
    g = (cx[0][[2, 3]].mean() - cx[0][[0, 1]].mean()) / h
    G = E_ / (2 * (1 + nu))
    sig = o.ip(F_STRESS)
    assert g == pytest.approx(1.0, rel=1e-6)
    np.testing.assert_allclose(sig[1], G * np.sin(g), rtol=2e-4)
    np.testing.assert_allclose(sig[0] - sig[3], 2 * G * (1 - np.cos(g)), rtol=2e-3)
    np.testing.assert_allclose(sig[5], 0.0, atol=1e-6 * G)
