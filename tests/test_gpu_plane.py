"""Plane-strain 4-node quad (ElQua4N2D) on the GPU against the FP64 oracle
(GPU), to the bounds of the other element types: FP64 nodal displacement
1e-9 normwise relative, IP fields 1e-7, time and increments exact; mixed
precision displacement 1e-5."""
import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from test_gpu_parity import MIXED_DISP_TOL, compare, compare_mixed, run_both
from test_oracle_plane import _quad_model
from test_oracle_solve import elastic_law

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt_method", [None, dnl.Explicit.PowerIteration])
def test_plate_necking_fp64(ctx, dt_method):
    m = samples.necking_plane(nx=8, ny=40, t_end=20e-6, pull=5.0, dt_method=dt_method, context=ctx)
    o = run_both(m, 20e-6)
    compare(m, o)
    assert np.max(m.getIntPointField(dnl.Field.plasticStrain)) > 1e-3  # yielded


def test_plate_necking_mixed(ctx):
    m = samples.necking_plane(nx=8, ny=40, t_end=20e-6, pull=5.0, precision="mixed", track_tensors=False,
                              context=ctx)
    o = run_both(m, 20e-6)
    e = compare_mixed(m, o, "mixed")
    assert e["disp"] <= MIXED_DISP_TOL


def test_plane_simple_shear_on_gpu(ctx):
    """The oracle's Jaumann simple-shear pin, run by the engine: one plane
    quad sheared to g = 0.2, s12 = G sin g."""
    h = 1.0
    xyz = np.array([[0, 0, 0], [h, 0, 0], [h, h, 0], [0, h, 0]], float)
    E_, nu, rho = 206e9, 0.3, 7830.0
    c = np.sqrt(E_ * (1 - nu) / (rho * (1 + nu) * (1 - 2 * nu)))
    dt = 0.9 * 1.981818700333 * (h / np.sqrt(2) / c) / 2
    V = 2e-4 * h / dt
    t_end = 0.2 * h / V
    m = _quad_model(xyz, [[0, 1, 2, 3]], elastic_law(), restrain_sets=[((1, 1, 0), [0, 1]), ((0, 1, 0), [2, 3])],
                    speed_sets=[((V, 0.0, 0.0), [2, 3])], t_end=t_end)
    m._ctx = ctx
    o = run_both(m, t_end)
    compare(m, o, fields=("Stress",))
    G = E_ / (2 * (1 + nu))
    c_ = m.getNodalCoordinates()
    g = (c_[0][[2, 3]].mean() - c_[0][[0, 1]].mean()) / h
    np.testing.assert_allclose(m.getIntPointField(dnl.Field.Stress)[1], G * np.sin(g), rtol=2e-4)
