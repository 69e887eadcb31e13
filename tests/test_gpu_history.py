"""History files taken on the GPU against the FP64 oracle's (GPU).

Bounds (FP64 vs FP64, as the field parity tests): the same rows (count and
time column to 1e-12 relative); per field, normwise relative to the
largest value of that field in the file, 1e-9 for nodal kinematics,
masses, kinetic energy and time steps, 1e-7 for IP-derived items.  The files hold the reference's
"%10.7E " text of the same rows."""
import os

import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import OracleJob

pytestmark = pytest.mark.gpu

F = dnl.Field
KIN_TOL, FIELD_TOL = 1e-9, 1e-7
KINEMATIC = {F.coordsX, F.coordsY, F.coordsZ, F.coords, F.dispX, F.dispY, F.dispZ, F.disp, F.speedX, F.speedY,
             F.speedZ, F.speed, F.dispIncX, F.dispIncY, F.dispIncZ, F.speedIncX, F.speedIncY, F.speedIncZ, F.mass,
             F.kineticEnergy, F.timeStep, F.realTimeStep}


def _hist(path, save, items, start=None):
    h = dnl.HistoryFile(os.path.basename(path))
    h.setFileName(str(path))
    if start is None:
        h.setSaveTime(save)
    else:
        h.setSaveTime(start, 1.0, save)
    for it in items:
        h.add(*it)
    return h


def _files(model, tmp, n_nodes, n_elems, save):
    ns = dnl.NodeSet("ns")
    model.add(ns, 1, n_nodes, max(1, n_nodes // 7))
    es = dnl.ElementSet("es")
    model.add(es, 1, n_elems, max(1, n_elems // 9))
    one = dnl.ElementSet("one")
    model.add(one, 1)
    nodal = [(ns, f) for f in (F.coordsX, F.coordsY, F.coords, F.dispY, F.speedY, F.speed, F.dispIncY,
                               F.speedIncY, F.mass)]
    node_ip = [(ns, f) for f in (F.T, F.plasticStrain, F.density, F.vonMises, F.StressYY, F.Stress, F.pressure,
                                 F.yieldStress, F.inelasticEnergy)]
    elem = [(es, 0, f) for f in (F.vonMises, F.plasticStrain, F.T, F.density, F.StressXX, F.StressXY, F.Stress,
                                 F.internalEnergy, F.pressure, F.gamma, F.gammaCumulate, F.plasticStrainRate)]
    elem += [(one, k, F.plasticStrain) for k in (1, 2, 3, 4)]
    glob = [(F.kineticEnergy,), (F.timeStep,), (F.realTimeStep,)]
    return [_hist(tmp / "nodal.plot", save, nodal), _hist(tmp / "nodeip.plot", save, node_ip),
            _hist(tmp / "elem.plot", save, elem), _hist(tmp / "glob.plot", save / 3, glob),
            _hist(tmp / "every.plot", 1e-15, glob[:2]), _hist(tmp / "late.plot", save, glob, start=save * 2.5)]


def _field_errors(h, g, r):
    """Normwise error per field: the columns of one field share its largest
    value in the file."""
    fields = [it[1] for it in h._items]
    out = {}
    for f in sorted(set(fields)):
        cols = [1 + k for k, ff in enumerate(fields) if ff == f]
        scale = np.abs(r[:, cols]).max()
        d = np.abs(g[:, cols] - r[:, cols]).max()
        out[f] = d / scale if scale > 0 else d
    return out


def _check(model, hists, o, ids):
    for h, fid in zip(hists, ids):
        g, r = h.getRows(), o.history_rows(fid)
        assert g.shape == r.shape, (h.getFileName(), g.shape, r.shape)
        np.testing.assert_allclose(g[:, 0], r[:, 0], rtol=1e-12, atol=0)
        errs = _field_errors(h, g, r)
        for f, err in errs.items():
            tol = KIN_TOL if f in KINEMATIC else FIELD_TOL
            assert err <= tol, (h.getFileName(), dnl.Field.getVtklabel(f), err)
        worst = max(errs.values())
        print(f"  {os.path.basename(h.getFileName())}: {len(g)} rows x {g.shape[1]} columns, worst {worst:.2e}")
        # the file: header, then the same rows as "%10.7E " text
        with open(h.getFileName()) as fh:
            lines = fh.read().splitlines()
        assert lines[0] == "#DynELA FEM Code v. 4.0 history file"
        assert lines[1] == "#plotted :" + "".join(n + " " for n in h.getItemNames())
        assert len(lines) == 2 + len(g)
        for line, row in zip(lines[2:], g):
            assert line == "".join("%10.7E " % v for v in row)


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_history_matches_oracle(ctx, tmp_path, kind):
    if kind == "axi":
        m = samples.taylor_axi(nr=5, nz=24, t_end=12e-6, context=ctx)
        nn, ne = 6 * 25, 5 * 24
    else:
        m = samples.taylor_3d(n=3, nz=10, t_end=10e-6, context=ctx)
        a = m.descriptors()
        nn, ne = a.coords.shape[1], a.conn.shape[0]
    t_end = m._solver.endTime
    hists = _files(m, tmp_path, nn, ne, t_end / 10)
    o = OracleJob(m.descriptors())
    ids = [o.add_history_file(m, h) for h in hists]
    for h in hists:
        m.add(h)
    assert o.solve(t_end) == 0
    m.solve()
    _check(m, hists, o, ids)


def test_history_over_many_launch_sequences(ctx, tmp_path):
    """A row every cycle for several hundred cycles (the device ring wraps)
    and a run in two solve calls: the rows are the oracle's."""
    m = samples.taylor_axi(nr=3, nz=8, t_end=80e-6, context=ctx)
    ns = dnl.NodeSet("top")
    m.add(ns, 33, 36)
    hists = [_hist(tmp_path / "every.plot", 1e-15, [(F.kineticEnergy,), (F.timeStep,), (ns, F.coordsY), (ns, F.T)]),
             _hist(tmp_path / "sparse.plot", 3e-6, [(ns, F.speedY)])]
    o = OracleJob(m.descriptors())
    ids = [o.add_history_file(m, h) for h in hists]
    for h in hists:
        m.add(h)
    assert o.solve(25e-6) == 0 and o.solve(80e-6) == 0
    m.solve(25e-6)
    m.solve(80e-6)
    assert len(hists[0].getRows()) > 256  # the device ring's capacity
    _check(m, hists, o, ids)


def test_history_rows_do_not_depend_on_launch_mode_or_batch(tmp_path):
    """Plain launches, captured sequences of several lengths, and a batch
    with a job without history files give bit-identical rows."""
    def run(cycles, partner):
        old = os.environ.get("DNLG_GRAPH_CYCLES")
        os.environ["DNLG_GRAPH_CYCLES"] = str(cycles)
        try:
            c = dnl.Context(0)
        finally:
            if old is None:
                del os.environ["DNLG_GRAPH_CYCLES"]
            else:
                os.environ["DNLG_GRAPH_CYCLES"] = old
        m = samples.taylor_axi(nr=4, nz=16, t_end=10e-6, context=c)
        ns = dnl.NodeSet("n")
        m.add(ns, 1, 85, 6)
        h = _hist(tmp_path / f"h{cycles}{partner}.plot", 0.7e-6,
                  [(ns, F.coordsY), (ns, F.vonMises), (F.kineticEnergy,), (F.timeStep,)])
        m.add(h)
        if partner:
            other = samples.taylor_3d(n=3, nz=8, t_end=10e-6, context=c)
            dnl.batch_solve([m, other])
        else:
            m.solve()
        return h.getRows()

    ref = run(16, False)
    for cycles, partner in ((0, False), (5, False), (16, True)):
        np.testing.assert_array_equal(run(cycles, partner), ref)


def test_history_mixed_precision(ctx, tmp_path):
    """Mixed precision: the rows of the FP64 oracle to the mixed path's
    bounds (1e-5 on kinematics), and the last row's element means equal to
    the mixed path's own field read-back."""
    m = samples.taylor_axi(nr=5, nz=24, t_end=12e-6, precision="mixed", track_tensors=False, context=ctx)
    es = dnl.ElementSet("es")
    m.add(es, 1, 120)
    ns = dnl.NodeSet("ns")
    m.add(ns, 1, 150, 11)
    h1 = _hist(tmp_path / "kin.plot", 1e-6, [(ns, F.coordsY), (ns, F.dispY), (F.kineticEnergy,)])
    h2 = _hist(tmp_path / "el.plot", 1e-15, [(es, 0, F.plasticStrain), (es, 0, F.gammaCumulate),
                                             (es, 0, F.vonMises), (es, 2, F.gamma)])
    o = OracleJob(m.descriptors())
    ids = [o.add_history_file(m, h) for h in (h1, h2)]
    m.add(h1)
    m.add(h2)
    assert o.solve(12e-6) == 0
    m.solve()
    g, r = h1.getRows(), o.history_rows(ids[0])
    assert g.shape == r.shape
    errs = _field_errors(h1, g, r)
    print({dnl.Field.getVtklabel(f): f"{e:.2e}" for f, e in errs.items()})
    assert max(errs.values()) <= 1e-5
    rows = h2.getRows()
    assert rows[-1, 0] == m.getCurrentTime()
    last = rows[-1, 1:].reshape(4, 120)
    ep = m.getIntPointField(F.plasticStrain)
    np.testing.assert_allclose(last[0], ep.mean(axis=1), rtol=1e-14)
    np.testing.assert_allclose(last[1], m.getIntPointField(F.gammaCumulate).mean(axis=1), rtol=1e-14)
    np.testing.assert_allclose(last[2], m.getIntPointField(F.vonMises).mean(axis=1), rtol=1e-12)
    np.testing.assert_array_equal(last[3], m.getIntPointField(F.gamma)[:, 1])


def test_history_file_of_the_taylor_sample(ctx, tmp_path):
    """Samples/Taylor/Taylor-Axi/Taylor.py's history set-up through the API:
    the kinetic-energy file starts at M v0^2 / 2 and decreases, the time-step
    file holds the time steps of the solve."""
    m = samples.taylor_axi(nr=5, nz=25, t_end=20e-6, context=ctx)
    speed = 227.0
    ke = dnl.HistoryFile("keHistory")
    ke.setFileName(str(tmp_path / "ke.plot"))
    ke.add(F.kineticEnergy)
    ke.setSaveTime(20e-6 / 50)
    m.add(ke)
    dt = dnl.HistoryFile("dtHistory")
    dt.setFileName(str(tmp_path / "dt.plot"))
    dt.add(F.timeStep)
    dt.setSaveTime(20e-6 / 50)
    m.add(dt)
    m.solve()
    rows = np.loadtxt(str(tmp_path / "ke.plot"), comments="#")
    assert rows.shape[0] in (50, 51) and rows.shape[1] == 2
    M = m.getNodalMasses().sum()
    assert rows[0, 0] == 0.0
    assert rows[0, 1] == pytest.approx(M * speed ** 2 / 2, rel=1e-7)
    assert rows[:, 1].max() <= rows[0, 1] * (1 + 1e-7) and rows[-1, 1] < 0.9 * rows[0, 1]
    d = np.loadtxt(str(tmp_path / "dt.plot"), comments="#")
    assert d[-1, 0] == pytest.approx(20e-6, rel=1e-7)
    assert np.all(d[:, 1] > 0) and np.all(np.diff(d[:, 0]) >= d[1:, 1] * (1 - 1e-6))
    assert d[-1, 1] == float("%10.7E" % m.getTimeStep())
