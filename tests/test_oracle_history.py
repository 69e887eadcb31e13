"""History files of the CPU oracle (no GPU): the save rule of
HistoryFile::save (HistoryFile.C:283-315) through its consequences, item
values against closed forms (free flight, uniform fields) and against the
field read-back, the IP-to-node extrapolation table against the shape
functions, the header names of the dnlPython HistoryFile; the node-field
read-back, the result files' default fields and the mesh transforms of
DynELA."""
import ctypes

import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import (F_DENSITY, F_PLASTIC_STRAIN, F_PRESSURE, F_STRESS, F_TEMPERATURE, F_VON_MISES, OracleJob,
                           extrapolation, load)
from test_oracle_solve import UNIT_HEX, _hex_model, elastic_law

QUAD, HEX = 3, 4
F = dnl.Field


def _tables(elem_type):
    N, D, w = np.zeros(64), np.zeros(192), np.zeros(8)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    nip = load().oracle_tables(elem_type, P(N), P(D), P(w))
    return N[: nip * nip].reshape(nip, nip)  # [ip][node]


@pytest.mark.parametrize("elem_type", [QUAD, HEX])
def test_extrapolation_inverts_the_shape_functions(elem_type):
    """integrationPointsToNode (ElQua4NAx.C:30-60, ElHex8N3D.C:35-90) maps IP
    values of any field in the element's shape-function space back to its
    nodal values: E N = I with N[ip][node] the shape functions at the Gauss
    points (themselves pinned against the closed forms)."""
    E = extrapolation(elem_type)
    N = _tables(elem_type)
    np.testing.assert_allclose(E @ N, np.eye(len(E)), atol=1e-14)
    np.testing.assert_allclose(E.sum(axis=1), 1.0, rtol=1e-14)


def _hist(name, save, *items, start=None):
    h = dnl.HistoryFile(name)
    h.setFileName(name + ".plot")
    if start is None:
        h.setSaveTime(save)
    else:
        h.setSaveTime(start, 1.0, save)
    for it in items:
        h.add(*it)
    return h


def _run(model, hists, t_end):
    o = OracleJob(model.descriptors())
    ids = [o.add_history_file(model, h) for h in hists]
    assert o.solve(t_end) == 0
    return o, [o.history_rows(i) for i in ids]


def test_save_rule_consequences():
    """Rows at the pre-loop pass and at the end of cycles: a tiny save
    interval takes every cycle (time column = the cycle times, timeStep
    column = their differences); an interval longer than the run takes the
    t = 0 row only; an interval D > dt takes exactly the first cycle time at
    or past each multiple of D; with a start time every cycle time t <= start
    takes a row."""
    t_end = 3e-6
    m = samples.taylor_axi(nr=3, nz=8, t_end=t_end)
    D = 0.4e-6
    every = _hist("every", 1e-15, (F.timeStep,))
    once = _hist("once", 1.0, (F.timeStep,))
    regular = _hist("regular", D, (F.timeStep,))
    started = _hist("started", 1e-6, (F.timeStep,), start=1.2e-6)
    o, (ra, rb, rc, rd) = _run(m, [every, once, regular, started], t_end)
    t, _, inc = o.scalars()
    times = ra[:, 0]
    assert len(ra) == inc + 1 and times[0] == 0.0 and times[-1] == t == t_end
    assert np.all(np.diff(times) > 0)
    np.testing.assert_allclose(np.diff(times), ra[1:, 1], rtol=1e-9)
    assert len(rb) == 1 and rb[0, 0] == 0.0
    assert ra[:, 1].max() < D
    expect = [0.0] + [times[np.argmax(times >= k * D)] for k in range(1, int(t_end / D) + 1) if times[-1] >= k * D]
    np.testing.assert_array_equal(rc[:, 0], expect)
    assert set(times[times <= 1.2e-6]) <= set(rd[:, 0])


def test_free_flight_items():
    """A free hex with speed v: coordinates x0 + v t, displacement v t, speed
    v, no acceleration, kinetic energy M v.v / 2 in every row.  dispInc is
    the reference's field0 increment at the end of the cycle, i.e. the
    previous cycle's v dt (0 in the first two rows)."""
    xyz = UNIT_HEX * 0.01
    v = np.array([3.0, -2.0, 1.5])
    m = _hex_model(xyz, [list(range(8))], elastic_law(), initial=tuple(v), t_end=2e-5, safety=0.5)
    ns = dnl.NodeSet("n7")
    m.add(ns, 7)
    items = [(ns, f) for f in (F.coordsX, F.coordsY, F.coordsZ, F.dispX, F.dispY, F.dispZ, F.speedX, F.speedY,
                               F.speedZ, F.speed, F.dispIncX, F.speedIncX, F.mass)]
    h = _hist("flight", 1e-15, *items, (F.kineticEnergy,), (F.timeStep,), (F.realTimeStep,))
    o, (r,) = _run(m, [h], 2e-5)
    t, dts = r[:, 0], r[:, -2]
    x0 = xyz[6]
    for d in range(3):
        np.testing.assert_allclose(r[:, 1 + d], x0[d] + v[d] * t, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(r[:, 4 + d], v[d] * t, rtol=1e-9, atol=1e-18)
        np.testing.assert_allclose(r[:, 7 + d], v[d], rtol=1e-10)
    np.testing.assert_allclose(r[:, 10], np.linalg.norm(v), rtol=1e-10)
    np.testing.assert_allclose(r[2:, 11], v[0] * dts[1:-1], rtol=1e-9)
    assert np.all(r[:2, 11] == 0.0)
    assert np.abs(r[:, 12]).max() < 1e-6 * np.abs(v).max() / dts.min()
    mass = o.mass()
    assert np.all(r[:, 13] == mass[6])
    np.testing.assert_allclose(r[:, 14], mass.sum() * v.dot(v) / 2, rtol=1e-10)
    np.testing.assert_array_equal(r[:, 15], r[:, 16])
    assert r[-1, 15] == o.scalars()[1]


def test_element_and_node_items_match_the_field_read_back():
    """Last row (t = t_end) against the IP fields read after the solve:
    element items are one IP's value or the IP mean (Element.C:1224-1238),
    node items the IP values extrapolated to the node and averaged over its
    elements (Node.C:244-300, 350-364); the element pressure is the IP's
    pressure member, equal to tr(Stress)/3 (the stress update adds it back).
    In the first row (t = 0, uniform T and density) the extrapolated nodal
    values are exactly the uniform values."""
    t_end = 6e-6
    m = samples.taylor_axi(nr=4, nz=12, t_end=t_end)
    a = m.descriptors()
    es = dnl.ElementSet("es")
    m.add(es, 1, 48)
    ns = dnl.NodeSet("ns")
    m.add(ns, 1, 65, 4)
    scalars = [F.vonMises, F.plasticStrain, F.T, F.density, F.pressure]
    elem_items = [(es, 0, f) for f in scalars] + [(es, k, F.plasticStrain) for k in (1, 2, 3, 4)]
    elem_items += [(es, 0, f) for f in (F.StressXX, F.StressXY, F.StressYY, F.StressZZ, F.Stress)]
    node_items = [(ns, f) for f in (F.T, F.plasticStrain, F.density, F.vonMises, F.StressXX, F.StressYX, F.Stress)]
    he = _hist("elements", 1e-6, *elem_items)
    hn = _hist("nodes", 1e-6, *node_items)
    o, (re_, rn) = _run(m, [he, hn], t_end)
    assert re_[-1, 0] == t_end
    ne, nip = o.NE, o.nip
    el = np.arange(48)
    last = re_[-1, 1:].reshape(-1, 48)
    ip = {f: o.ip(k) for f, k in ((F.vonMises, F_VON_MISES), (F.plasticStrain, F_PLASTIC_STRAIN),
                                  (F.T, F_TEMPERATURE), (F.density, F_DENSITY), (F.pressure, F_PRESSURE))}
    sig = o.ip(F_STRESS)
    for k, f in enumerate(scalars):
        tol = 1e-12 * np.abs(sig).max() if f == F.pressure else 1e-13 * np.abs(ip[f]).max()
        np.testing.assert_allclose(last[k], ip[f][el].mean(axis=1), rtol=1e-13, atol=tol)
    for k in range(4):
        np.testing.assert_array_equal(last[5 + k], ip[F.plasticStrain][el, k])
    for k, c in enumerate((0, 1, 3, 5)):
        np.testing.assert_allclose(last[9 + k], sig[c][el].mean(axis=1), rtol=1e-13, atol=1e-13 * np.abs(sig).max())
    norm = np.sqrt(sig[0] ** 2 + sig[3] ** 2 + sig[5] ** 2 + 2 * (sig[1] ** 2 + sig[2] ** 2 + sig[4] ** 2))
    np.testing.assert_allclose(last[13], norm[el].mean(axis=1), rtol=1e-13)
    # node items: sum over the node's elements of E[local] . IP values, over the count
    E = extrapolation(QUAD)
    nodes = ns.numbers() - 1  # numbering from 1, nodes in number order
    conn = a.conn

    def at_node(vals, n):
        acc, cnt = 0.0, 0
        for e in range(ne):
            loc = np.nonzero(conn[e] == n)[0]
            if len(loc):
                acc = acc + E[loc[0]] @ vals[e].T
                cnt += 1
        return acc / cnt

    lastn = rn[-1, 1:].reshape(-1, len(nodes))
    for k, f in enumerate((F.T, F.plasticStrain, F.density, F.vonMises)):
        want = np.array([at_node(ip[f], n) for n in nodes])
        np.testing.assert_allclose(lastn[k], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    sxx = np.array([at_node(sig[0], n) for n in nodes])
    np.testing.assert_allclose(lastn[4], sxx, rtol=1e-12, atol=1e-12 * np.abs(sxx).max())
    np.testing.assert_allclose(lastn[5], np.array([at_node(sig[1], n) for n in nodes]), rtol=1e-12,
                               atol=1e-12 * np.abs(sxx).max())
    ten = np.array([[at_node(sig[c], n) for c in range(6)] for n in nodes]).T
    tn = np.sqrt(ten[0] ** 2 + ten[3] ** 2 + ten[5] ** 2 + 2 * (ten[1] ** 2 + ten[2] ** 2 + ten[4] ** 2))
    np.testing.assert_allclose(lastn[6], tn, rtol=1e-12)
    first = rn[0, 1:].reshape(-1, len(nodes))
    assert rn[0, 0] == 0.0
    np.testing.assert_allclose(first[0], samples.COPPER["T0"], rtol=1e-14)
    np.testing.assert_allclose(first[2], samples.COPPER["density"], rtol=1e-14)


def test_history_item_names():
    """HistoryFile::add names (HistoryFile.C:186-280): label_n<number> and
    label[_intPt]_e<number> for sets of several members, the bare label for
    one member; yieldStress is labelled "yield"."""
    many, one = dnl.NodeSet("many"), dnl.NodeSet("one")
    many._numbers.extend([3, 12])
    one._numbers.append(5)
    es = dnl.ElementSet("es")
    es._numbers.extend([7, 9])
    e1 = dnl.ElementSet("e1")
    e1._numbers.append(4)
    h = dnl.HistoryFile("h")
    h.add(many, F.coordsX)
    h.add(one, F.yieldStress)
    h.add(es, 2, F.plasticStrain)
    h.add(es, 0, F.StressXY)
    h.add(e1, 0, F.T)
    h.add(F.kineticEnergy)
    assert h.getItemNames() == ["coordsX_n3", "coordsX_n12", "yield", "plasticStrain_2_e7", "plasticStrain_2_e9",
                                "StressXY_e7", "StressXY_e9", "T", "kineticEnergy"]
    assert F.coordsX == 1 and F.density == 20 and F.yieldStress == 34 and F.Stress == 75 and F.ENDFIELDS == 85
    assert F.getField("yield") == F.yieldStress


def test_vtk_default_fields():
    """VtkInterface::initFields with the configuration's VtkFields: labels
    that name no field (temperature, displacement, displacementIncrement) are
    left out."""
    v = dnl.VtkInterface()
    assert v._fields == [F.Stress, F.Strain, F.PlasticStrain, F.vonMises, F.yieldStress, F.pressure, F.plasticStrain,
                         F.plasticStrainRate, F.gamma, F.gammaCumulate, F.speed]
    v.addField(F.T)
    v.removeField(F.Strain)
    assert v.getNumberOfFields() == 11 and v.existField(F.T) == 10 and v.existField(F.Strain) == -1


def test_geometry_transforms():
    """DynELA::translate / scale / rotate (DynELA.C:542-661) on the node
    coordinates: a quarter turn about Z takes x to y; the reference's "X" and
    "Y" matrices turn the other way (y -> -z, z -> -x); about an axis the
    right-handed rotation; on a node set only its nodes move."""
    def model():
        m = dnl.DynELA("t")
        m.createNodes([1, 2, 3], [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
        return m

    def xyz(m):
        return m._node_tables()[1]

    m = model()
    m.rotate("Z", 90.0)
    np.testing.assert_allclose(xyz(m), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]], atol=1e-15)
    m = model()
    m.rotate("X", 90.0)
    np.testing.assert_allclose(xyz(m), [[1, 0, 0], [0, 0, -1], [0, 1, 0]], atol=1e-15)
    m = model()
    m.rotate("Y", 90.0)
    np.testing.assert_allclose(xyz(m), [[0, 0, 1], [0, 1, 0], [-1, 0, 0]], atol=1e-15)
    m = model()
    m.rotate(dnl.Vec3D(1, 1, 1), 120.0)  # the cyclic permutation x -> y -> z
    np.testing.assert_allclose(xyz(m), [[0, 1, 0], [0, 0, 1], [1, 0, 0]], atol=1e-15)
    m = model()
    ns = dnl.NodeSet("2")
    m.add(ns, 2)
    m.translate(dnl.Vec3D(1, 2, 3), ns)
    m.scale(2.0)
    m.scale(dnl.Vec3D(1, 0.5, 1), ns)
    np.testing.assert_allclose(xyz(m), [[2, 0, 0], [2, 3, 6], [0, 0, 2]], atol=1e-15)
    lo, hi = m.getGlobalBox()
    assert (lo(0), lo(1), lo(2), hi(0), hi(1), hi(2)) == (0, 0, 0, 2, 3, 6)


def test_node_field_read_back_pins():
    """Node::fieldScalar of every node: uniform IP fields extrapolate to the
    same value at every node (weights sum to 1, divided by the element
    count), nodal fields are the node arrays, unknown fields are 0."""
    m = samples.taylor_axi(nr=3, nz=8, t_end=1e-6)
    o = OracleJob(m.descriptors())
    np.testing.assert_allclose(o.node_field(F.T), samples.COPPER["T0"], rtol=1e-14)
    np.testing.assert_allclose(o.node_field(F.density), samples.COPPER["density"], rtol=1e-14)
    assert o.solve(1e-6) == 0
    c, d, v = o.nodes()
    np.testing.assert_array_equal(o.node_field(F.coordsY), c[1])
    np.testing.assert_array_equal(o.node_field(F.speedY), v[1])
    np.testing.assert_array_equal(o.node_field(F.mass), o.mass())
    np.testing.assert_allclose(o.node_field(F.disp), np.linalg.norm(d, axis=0), rtol=1e-15)
    assert np.all(o.node_field(F.inelasticEnergy) == 0.0) and np.all(o.node_field(F.timeStep) == 0.0)


def test_svg_interface_is_accepted_but_writes_nothing():
    """SvgInterface (not provided): the samples' view settings are accepted,
    write() warns instead of writing."""
    svg = dnl.SvgInterface("SVG")
    svg.setTitleDisplay(False)
    svg.setLegendPosition(350, 150)
    svg.rotate(dnl.Vec3D(1, 0, 0), -90)
    with pytest.warns(RuntimeWarning, match="not provided"):
        svg.write("temperature.svg", F.T)
    with pytest.raises(AttributeError):
        svg.drawSomething()


def test_element_view_and_estimated_end_without_a_job(capsys):
    """getElementByNum (DynELA.C:144-156): the element's node numbers in
    connectivity order, found by element number whatever the creation
    order; displayEstimatedEnd prints nothing before any time has passed."""
    m = dnl.DynELA("v")
    m.createNodes(np.array([10, 20, 30, 40, 50, 60]),
                  np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0], [1, 1, 0], [2, 1, 0]], float))
    m.setDefaultElement(dnl.Element.ElQua4N2D)
    m.createElements(np.array([7, 3]), np.array([[20, 30, 60, 50], [10, 20, 50, 40]]))
    es = dnl.ElementSet("all")
    m.add(es, 3, 7, 4)
    mat = dnl.Material("m")
    mat.youngModulus, mat.poissonRatio, mat.density, mat.heatCapacity = 1e9, 0.3, 1000.0, 400.0
    law = dnl.ElastoplasticLaw("p")
    law.setParameters(1e6, 0.0, 1.0)
    mat.setHardeningLaw(law)
    m.add(mat, es)
    s = dnl.Explicit("s")
    s.setTimes(0.0, 1e-6)
    m.add(s)
    e = m.getElementByNum(7)
    assert e.nodes == [20, 30, 60, 50] and e.getNumberOfNodes() == 4 and e.getNumberOfIntegrationPoints() == 4
    assert m.getElementByNum(3).nodes == [10, 20, 50, 40]
    with pytest.raises(dnl.DnlgError):
        m.getElementByNum(5)
    m.displayEstimatedEnd()
    assert capsys.readouterr().out == ""


def test_set_increments_does_not_bound_the_explicit_loop():
    """Solver::setIncrements is accepted; the explicit loop runs on time
    only (Explicit.C:184), so the solve's descriptors do not change."""
    m1 = samples.taylor_axi(nr=3, nz=8, t_end=2e-6)
    m2 = samples.taylor_axi(nr=3, nz=8, t_end=2e-6)
    m2._solver.setIncrements(0, 3)
    a, b = m1.descriptors().params, m2.descriptors().params
    assert bytes(a) == bytes(b)


def test_set_operations_and_material_constants():
    """NodeSet / ElementSet size, getNumber, intersect, substract
    (NodeSet.C:87-160) and the Material constants of Material.h."""
    a, b = dnl.NodeSet("a"), dnl.NodeSet("b")
    a._numbers.extend([5, 1, 9, 3])
    b._numbers.extend([9, 5, 7])
    a.intersect(b)
    assert a.size() == 2 and [a.getNumber(i) for i in range(2)] == [5, 9]
    a._numbers.extend([4, 7])
    a.substract(b)
    assert a._numbers == [4]
    a.clear()
    assert a.size() == 0
    mat = dnl.Material("m")
    mat.youngModulus, mat.poissonRatio, mat.density, mat.heatCapacity, mat.taylorQuinney = 200e9, 0.3, 7800.0, 460.0, 0.9
    G = mat.getShearModulus()
    assert mat.getTwoShearModulus() == pytest.approx(2 * G, rel=1e-15)
    lam = mat.getLambda()
    assert mat.getBulkModulus() == pytest.approx(lam + 2 * G / 3, rel=1e-14)
    assert mat.getHeatFraction() == pytest.approx(0.9 / (7800.0 * 460.0), rel=1e-15)
    assert not mat.isYield()
    mat.setHardeningLaw(dnl.ElasticLaw("e"))
    assert not mat.isYield()
    mat.setHardeningLaw(dnl.JohnsonCookLaw("j"))
    assert mat.isYield()
