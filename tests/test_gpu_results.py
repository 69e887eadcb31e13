"""Node fields and result (VTK) files on the GPU (GPU).

Node fields (Node::fieldScalar of every node) against the FP64 oracle's,
normwise relative per field: 1e-9 for nodal kinematics, 1e-7 for IP fields
extrapolated to the nodes.  Result files: written where the reference writes
them (initial configuration, every cycle that reaches the next save time,
final state), in VtkInterface's layout, with the values of the engine's
history rows at the same cycle (the files print 6 significant digits); a
run that writes them ends in the same state, bit for bit, as one that does
not."""
import os

import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import OracleJob

pytestmark = pytest.mark.gpu

F = dnl.Field


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_node_fields_match_oracle(ctx, kind):
    if kind == "axi":
        m = samples.taylor_axi(nr=5, nz=24, t_end=12e-6, context=ctx)
    else:
        m = samples.taylor_3d(n=3, nz=10, t_end=10e-6, context=ctx)
    o = OracleJob(m.descriptors())
    assert o.solve(m._solver.endTime) == 0
    m.solve()
    kin = (F.coordsX, F.coordsY, F.coords, F.dispY, F.disp, F.speedY, F.speed, F.dispIncY, F.speedIncY, F.mass)
    ipf = (F.T, F.plasticStrain, F.plasticStrainRate, F.density, F.vonMises, F.pressure, F.yieldStress,
           F.gammaCumulate, F.internalEnergy, F.StressXX, F.StressYZ, F.Stress, F.Strain, F.PlasticStrainZZ)
    for f, tol in [(f, 1e-9) for f in kin] + [(f, 1e-7) for f in ipf]:
        g, r = m.getNodalField(f), o.node_field(f)
        scale = np.abs(r).max()
        err = np.abs(g - r).max() / scale if scale > 0 else np.abs(g).max()
        assert err <= tol, (F.getVtklabel(f), err)
    assert np.all(m.getNodalField(F.inelasticEnergy) == 0.0)
    lo, hi = m.getNodalValuesRange(F.plasticStrain)
    assert lo == m.getNodalField(F.plasticStrain).min() and hi > 0.1


def _parse_vtk(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    assert lines[:6] == ["# vtk DataFile Version 4.0", "DynELA FEM Code v.4.0", "ASCII", "DATASET UNSTRUCTURED_GRID",
                         "FIELD FieldData 1", "TIME 1 1 double"]
    out = {"time": float(lines[6])}
    k = 7
    n = int(lines[k].split()[1])
    assert lines[k] == f"POINTS {n} float"
    out["points"] = np.array([[float(v) for v in lines[k + 1 + i].split()] for i in range(n)])
    k += n + 2
    ne, tot = (int(v) for v in lines[k].split()[1:])
    out["cells"] = [lines[k + 1 + i] for i in range(ne)]
    k += ne + 2
    assert lines[k] == f"CELL_TYPES {ne}"
    out["types"] = {lines[k + 1 + i] for i in range(ne)}
    k += ne + 2
    assert lines[k] == f"POINT_DATA {n}"
    k += 1
    fields = {}
    while k < len(lines) and lines[k]:
        kind, name, _ = lines[k].split()
        k += 1
        if kind == "SCALARS":
            assert lines[k] == "LOOKUP_TABLE default"
            k += 1
        fields[name] = (kind, np.array([[float(v) for v in lines[k + i].split()] for i in range(n)]))
        k += n
    out["fields"] = fields
    return out


def test_result_files(ctx, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    t_end = 12e-6

    def build(with_results):
        m = samples.taylor_axi(nr=5, nz=24, t_end=t_end, context=ctx)
        m.name = "Taylor"
        ns = dnl.NodeSet("all")
        m.add(ns, 1, 150)
        h = dnl.HistoryFile("every")
        h.setFileName(str(tmp_path / f"every{int(with_results)}.plot"))
        h.setSaveTime(1e-15)
        for f in (F.coordsX, F.coordsY, F.speedY, F.vonMises, F.plasticStrain, F.StressYY):
            h.add(ns, f)
        m.add(h)
        if with_results:
            m.setSaveTimes(0.0, t_end, t_end / 5)
        return m, h

    m, h = build(True)
    m.solve()
    files = sorted(p for p in os.listdir(tmp_path) if p.endswith(".vtk"))
    # the initial configuration, one per save time reached, the final state
    assert files[0] == "Taylor000.vtk" and 6 <= len(files) <= 7, files
    rows = h.getRows()
    first = _parse_vtk(tmp_path / files[0])
    assert first["time"] == 0.0
    c0 = m.descriptors().conn[0]
    assert first["types"] == {"9"} and first["cells"][0] == "4 " + "".join(f"{v} " for v in c0)
    assert list(first["fields"]) == ["Stress", "Strain", "PlasticStrain", "vonMises", "yield", "pressure",
                                     "plasticStrain", "plasticStrainRate", "gamma", "gammaCumulate", "speed"]
    assert first["fields"]["Stress"][0] == "TENSORS" and first["fields"]["speed"][0] == "VECTORS"
    np.testing.assert_allclose(first["fields"]["speed"][1][:, 1], -227.0)
    assert np.all(first["fields"]["vonMises"][1] == 0.0)
    # files written during the solve: the history row of the same cycle
    for name in files[1:-1]:
        v = _parse_vtk(tmp_path / name)
        r = rows[np.argmin(np.abs(rows[:, 0] - v["time"]))]
        assert abs(r[0] - v["time"]) <= 1e-5 * v["time"]
        cols = r[1:].reshape(6, 150)
        # the same doubles print as C++ streams print them ("%g")
        with open(tmp_path / name) as fh:
            text = fh.read().splitlines()
        assert text[6] == "%g" % r[0]
        assert text[8:158] == ["%g %g %g" % (x, y, 0.0) for x, y in zip(cols[0], cols[1])]
        for got, want in ((v["points"][:, 0], cols[0]), (v["points"][:, 1], cols[1]),
                          (v["fields"]["speed"][1][:, 1], cols[2]), (v["fields"]["vonMises"][1][:, 0], cols[3]),
                          (v["fields"]["plasticStrain"][1][:, 0], cols[4]), (v["fields"]["Stress"][1][:, 4], cols[5])):
            np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    last = _parse_vtk(tmp_path / files[-1])
    assert last["time"] == pytest.approx(t_end, rel=1e-5)
    np.testing.assert_allclose(last["fields"]["speed"][1].T, m.getNodalSpeeds()[:, :], rtol=1e-5,
                               atol=1e-5 * np.abs(m.getNodalSpeeds()).max())
    # no effect on the solve
    m2, h2 = build(False)
    m2.solve()
    np.testing.assert_array_equal(m.getNodalDisplacements(), m2.getNodalDisplacements())
    np.testing.assert_array_equal(rows, h2.getRows())


def test_element_view_values_and_estimated_end(ctx, capsys):
    """getElementByNum(n).getIntPointValue(field, ip) is the model's
    integration-point field at that element and IP; displayEstimatedEnd
    prints the reference's two lines once time has passed."""
    import re
    m = samples.taylor_axi(nr=4, nz=12, t_end=6e-6, context=ctx)
    m.solve(3e-6)
    ep = m.getIntPointField(dnl.Field.plasticStrain)
    nums = np.sort(np.concatenate(m._elem_num))
    for k in (0, len(nums) // 2, len(nums) - 1):
        e = m.getElementByNum(int(nums[k]))
        for ip in range(4):
            assert e.getIntPointValue(dnl.Field.plasticStrain, ip) == ep[k, ip]
    m.displayEstimatedEnd()
    out = capsys.readouterr().out.splitlines()
    assert re.fullmatch(r"Current computational time \d\d:\d\d:\d\d", out[0])
    assert re.fullmatch(r"Estimated end of computation in \d\d:\d\d:\d\d", out[1])
