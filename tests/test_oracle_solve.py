"""Whole-solve pins of the CPU oracle (no GPU): closed forms and the
reference's own published results for this path."""
import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import (F_INTERNAL_ENERGY, F_PLASTIC_STRAIN, F_STRESS, F_TEMPERATURE, OracleJob)


def _hex_model(xyz, conn, mat_law, speed_sets=(), restrain_sets=(), initial=None, t_end=1e-5, freq=10,
               safety=0.9):
    m = dnl.DynELA("test")
    m.createNodes(np.arange(1, len(xyz) + 1), xyz)
    m.setDefaultElement(dnl.Element.ElHex8N3D)
    m.createElements(np.arange(1, len(conn) + 1), np.asarray(conn) + 1)
    es = dnl.ElementSet("all")
    m.add(es, 1, len(conn))
    mat = dnl.Material("m")
    mat.youngModulus, mat.poissonRatio, mat.density = 206e9, 0.3, 7830.0
    mat.heatCapacity, mat.taylorQuinney, mat.T0 = 460.0, 0.9, 20.0
    mat.setHardeningLaw(mat_law)
    m.add(mat, es)
    for vals, nodes in restrain_sets:
        b = dnl.BoundaryRestrain("r")
        b.setValue(*vals)
        ns = dnl.NodeSet("n")
        ns._numbers.extend(int(n) + 1 for n in nodes)
        m.attachConstantBC(b, ns)
    for vals, nodes in speed_sets:
        b = dnl.BoundarySpeed("s")
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
    s.setComputeTimeStepFrequency(freq)
    s.setTimeStepSafetyFactor(safety)
    m.add(s)
    return m


UNIT_HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def elastic_law():
    law = dnl.ElastoplasticLaw("elastic")
    law.setParameters(1e30, 0.0, 1.0)  # yield never reached
    return law


def test_free_flight_is_rigid_translation():
    """No boundary, uniform initial speed: no strain, x(t) = x0 + v t.  For a
    single free hex the Courant length V/A_max underestimates the element's
    highest eigenfrequency, and at the default safety 0.9 the rounding noise
    of ln(lambda ~ 1) grows from step to step; a safety of 0.5 keeps the
    scheme inside its stability limit."""
    xyz = UNIT_HEX * 0.01
    v = (3.0, -2.0, 1.5)
    m = _hex_model(xyz, [list(range(8))], elastic_law(), initial=v, t_end=2e-5, safety=0.5)
    o = OracleJob(m.descriptors())
    assert o.solve(2e-5) == 0
    t, dt, inc = o.scalars()
    c, d, sp = o.nodes()
    np.testing.assert_allclose(d, np.array(v)[:, None] * t * np.ones((3, 8)), rtol=1e-12)
    np.testing.assert_allclose(sp, np.array(v)[:, None] * np.ones((3, 8)), rtol=1e-10)
    assert np.abs(o.ip(F_STRESS)).max() < 1e-2  # Pa: rounding of ln(lambda ~ 1) only


def test_lumped_mass_equals_rho_volume():
    m = samples.taylor_3d(n=6, nz=8)
    a = m.descriptors()
    o = OracleJob(a)
    o.solve(1e-9)
    x, conn = a.coords, a.conn
    # O-grid prism: volume = sum of cross-section quad areas x length
    nl = conn.shape[0] // 8
    q = conn[:nl, :4]
    xs, zs = x[0][q], x[2][q]
    area = 0.5 * np.abs(np.sum(xs * np.roll(zs, -1, 1) - np.roll(xs, -1, 1) * zs, axis=1)).sum()
    vol = area * 32.4e-3
    assert o.mass().sum() == pytest.approx(8930.0 * vol, rel=1e-12)
    ax = samples.taylor_axi(nr=4, nz=6)
    oa = OracleJob(ax.descriptors())
    oa.solve(1e-9)
    assert oa.mass().sum() == pytest.approx(8930.0 * np.pi * 3.2e-3 ** 2 * 32.4e-3, rel=1e-12)


def test_taylor_axi_energy_balance():
    """Impact on a frictionless rigid wall: the wall does no work, so kinetic
    + internal energy stays at the initial kinetic energy of the free DOFs,
    minus the small high-frequency dissipation of the scheme (rho_b 0.8182)."""
    m = samples.taylor_axi(nr=5, nz=25, t_end=30e-6)
    a = m.descriptors()
    o = OracleJob(a)
    o.solve(1e-12)
    ke_free = o.energies()[0]
    o.solve(30e-6)
    ke, ie, wp, tot = o.energies()
    assert ke_free > 0 and ie > 0.2 * ke_free
    assert tot / ke_free == pytest.approx(1.0, abs=0.02)
    assert tot <= ke_free * 1.001  # no energy created beyond the O(dt) discretisation error
    assert wp <= ie


def test_simple_shear_jaumann_closed_form():
    """One hex in homogeneous simple shear, F = I + gamma e1 (x) e2, every node
    prescribed: bottom fixed, top moving at speed V along x.  Rotating the
    stress by the polar rotation of each increment (Element.C:1146-1161)
    integrates the Jaumann (spin) rate, so the elastic solution is the
    Jaumann one of Samples/Element/Shear/Elastic/Analytic:
      s12 = G sin(gamma),  s11 - s22 = 2 G (1 - cos(gamma)),
    and not the Green-Naghdi one (Dienes 1979), which differs by 3 % at
    gamma = 1."""
    h = 1.0
    xyz = UNIT_HEX * h
    bottom, top = [0, 1, 4, 5], [2, 3, 6, 7]  # y = 0 / y = h
    gamma_end = 1.0
    # shear steps of about 2e-4: the time step is the elastic Courant step
    E, nu, rho = 206e9, 0.3, 7830.0
    c = np.sqrt(E * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))
    dt = 0.9 * 1.981818700333 * (h / c) / 2
    V = 2e-4 * h / dt
    t_end = gamma_end * h / V
    m = _hex_model(xyz, [list(range(8))], elastic_law(),
                   restrain_sets=[((1, 1, 1), bottom), ((0, 1, 1), top)],
                   speed_sets=[((V, 0.0, 0.0), top)], t_end=t_end)
    o = OracleJob(m.descriptors())
    assert o.solve(t_end) == 0
    cx, _, _ = o.nodes()
    g = (cx[0][top].mean() - cx[0][bottom].mean()) / h
    G = E / (2 * (1 + nu))
    sig = o.ip(F_STRESS)  # [6, 1, 8]
    assert g == pytest.approx(gamma_end, rel=1e-6)
    np.testing.assert_allclose(sig[1], G * np.sin(g), rtol=2e-4)
    np.testing.assert_allclose(sig[0] - sig[3], 2 * G * (1 - np.cos(g)), rtol=2e-3)
    np.testing.assert_allclose(sig[5], 0.0, atol=1e-6 * G)
    phi = np.arctan(g / 2)
    s12_gn = 2 * G * np.cos(2 * phi) * (2 * phi - 2 * np.tan(2 * phi) * np.log(np.cos(phi)) - np.tan(phi))
    assert abs(sig[1].mean() - s12_gn) > 0.02 * abs(s12_gn)


@pytest.mark.slow
def test_taylor_axi_published_results():
    """Samples/Taylor/Taylor-Axi (Taylor.py:12-34): 4340 steel, 5 x 50 quads,
    287 m/s, 80 us.  Published DynELA results (Documentation/Samples-
    Impacts.tex:106-107): final length 26.52 mm, impact-face diameter
    11.15 mm, centre-element plastic strain 1.78 and temperature 582.34 C.
    The sample uses the power-iteration time step; the Courant step used
    here changes only the time grid; the results agree to the published
    digits (0.1 % on lengths, under 1 % on the centre-element fields)."""
    steel = dict(young=206000.0, poisson=0.3, density=7.83e-09, heatCapacity=4.6e+08, taylorQuinney=0.9,
                 A=806.0, B=614.0, C=0.0089, n=0.168, m=1.1, depsp0=1.0, Tm=1540.0, T0=20.0)
    m = samples.taylor_axi(nr=5, nz=50, radius=3.2, length=32.4, speed=287000.0, t_end=80e-6, material=steel)
    o = OracleJob(m.descriptors(), threads=4)
    assert o.solve(80e-6) == 0
    c, _, _ = o.nodes()
    L_f = c[1].max()
    D_f = 2 * c[0][:6].max()
    ep = o.ip(F_PLASTIC_STRAIN)[0].mean()
    T = o.ip(F_TEMPERATURE)[0].mean()
    print(f"L_f={L_f:.3f} mm D_f={D_f:.3f} mm ep={ep:.3f} T={T:.1f} C")
    assert L_f == pytest.approx(26.52, rel=1e-3)
    assert D_f == pytest.approx(11.15, rel=1e-3)
    assert ep == pytest.approx(1.78, rel=1e-2)
    assert T == pytest.approx(582.34, rel=2e-3)
