"""Tabular hardening law (TabularLaw.C, DiscreteFunction.C:188-308) in the
CPU oracle (no GPU): the table interpolation against the dnlPython
DiscreteFunction's, a table sampled from the power law reproducing the
power-law solve, and the out-of-table error."""
import ctypes

import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import _lib, samples
from oracle.bridge import F_PLASTIC_STRAIN, F_STRESS, OracleJob, load


def _table_material(xs, ys):
    f = dnl.DiscreteFunction("table")
    for x, y in zip(xs, ys):
        f.add(x, y)
    law = dnl.TabularLaw("tab")
    law.setFunction(f)
    return law, f


def test_tabular_value_and_slope():
    """Interpolation at and between the points, the slope of the segment
    ending at the first point at or past x, out-of-range evaluation."""
    xs, ys = [0.0, 0.1, 0.3, 1.0], [100.0, 150.0, 170.0, 240.0]
    law, f = _table_material(xs[::-1], ys[::-1])  # points given in any order
    mat = dnl.Material("m")
    mat.setHardeningLaw(law)
    d = mat.descriptor()
    lib = load()
    for x, v, s in ((0.0, 100.0, 500.0), (0.05, 125.0, 500.0), (0.1, 150.0, 500.0), (0.2, 160.0, 100.0),
                    (1.0, 240.0, 100.0)):
        assert lib.oracle_law_yield(ctypes.byref(d), x, 0.0, 20.0) == pytest.approx(v, rel=1e-15) == f.getValue(x)
        assert lib.oracle_law_dyield(ctypes.byref(d), x, 0.0, 20.0, 1e-6) == pytest.approx(s, rel=1e-13)
        assert f.getSlope(x) == pytest.approx(s, rel=1e-13)
    assert np.isnan(lib.oracle_law_yield(ctypes.byref(d), 1.5, 0.0, 20.0))
    with pytest.raises(dnl.DnlgError):
        f.getValue(-0.1)


def test_tabular_power_law_table_reproduces_the_power_law():
    """A dense table through A + B ep^n (with a linear first segment from
    ep = 0) gives the power-law solve to the table's interpolation error."""
    A, B, n = 120e6, 400e6, 0.5

    def run(law):
        m = samples.taylor_axi(nr=4, nz=20, t_end=8e-6)
        m._material.setHardeningLaw(law)
        m._invalidate()
        o = OracleJob(m.descriptors(), threads=4)
        assert o.solve(8e-6) == 0
        return o

    power = dnl.ElastoplasticLaw("p")
    power.setParameters(A, B, n)
    xs = np.concatenate([[0.0], np.geomspace(1e-9, 5.0, 4000)])
    law, _ = _table_material(xs, A + B * xs ** n)
    a, b = run(power), run(law)
    ea, eb = a.ip(F_PLASTIC_STRAIN), b.ip(F_PLASTIC_STRAIN)
    assert ea.max() > 0.05
    assert np.abs(ea - eb).max() <= 2e-3 * ea.max()
    sa, sb = a.ip(F_STRESS), b.ip(F_STRESS)
    assert np.abs(sa - sb).max() <= 2e-3 * np.abs(sa).max()


def test_tabular_out_of_table_stops_the_solve():
    law, _ = _table_material([0.0, 0.01], [120e6, 125e6])
    m = samples.taylor_axi(nr=4, nz=20, t_end=8e-6)
    m._material.setHardeningLaw(law)
    m._invalidate()
    o = OracleJob(m.descriptors())
    assert o.solve(8e-6) == _lib.E_TABLE_RANGE
