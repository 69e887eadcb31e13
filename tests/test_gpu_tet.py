"""The 4-node tetrahedron (ElTet4N3D.C, one integration point) on the GPU
against the FP64 oracle (GPU), to the FP64 bounds of the other element
types: nodal displacement 1e-9 normwise relative, IP fields 1e-7, time and
increments exact.  The reference runs this element only under a
power-iteration time step (its characteristic length is 0) and the engine
runs it in FP64 only; both refusals are checked."""
import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from test_gpu_parity import compare, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt_method", [dnl.Explicit.PowerIteration, dnl.Explicit.PowerIterationUnder])
def test_tet_impact_against_oracle(ctx, dt_method):
    m = samples.taylor_tet(n=3, nz=12, t_end=10e-6, freq=1, dt_method=dt_method, context=ctx)
    o = run_both(m, 10e-6)
    compare(m, o)
    assert np.max(m.getIntPointField(dnl.Field.plasticStrain)) > 0.1


def test_tet_node_fields_and_masses(ctx):
    """Lumped masses, and node values of IP fields through the reference's
    extrapolation vectors (ElTet4N3D.C:33-60), against the oracle."""
    from oracle.bridge import OracleJob
    m = samples.taylor_tet(n=2, nz=8, t_end=4e-6, context=ctx)
    o = OracleJob(m.descriptors())
    assert o.solve(4e-6) == 0
    m.solve()
    np.testing.assert_allclose(m.getNodalMasses(), o.mass(), rtol=1e-13)
    for f in (dnl.Field.plasticStrain, dnl.Field.T, dnl.Field.vonMises):
        g = m.getNodalField(f)
        r = o.node_field(f)
        err = np.abs(g - r).max() / np.abs(r).max()
        print(f"node field {f}: {err:.2e}")
        assert err <= 1e-7


def test_tet_in_a_batch_with_a_hex_job(ctx):
    """A tetrahedral job and a hexahedral job in one batch: each is
    bit-identical to its solo run."""
    pi = dnl.Explicit.PowerIteration
    a = samples.taylor_tet(n=3, nz=12, t_end=6e-6, context=ctx)
    b = samples.taylor_3d(n=2, nz=10, t_end=6e-6, dt_method=pi, context=ctx)
    sa = samples.taylor_tet(n=3, nz=12, t_end=6e-6, context=ctx)
    sb = samples.taylor_3d(n=2, nz=10, t_end=6e-6, dt_method=pi, context=ctx)
    dnl.batch_solve([a, b])
    sa.solve()
    sb.solve()
    np.testing.assert_array_equal(a.getNodalDisplacements(), sa.getNodalDisplacements())
    np.testing.assert_array_equal(b.getNodalDisplacements(), sb.getNodalDisplacements())


@pytest.mark.parametrize("what", ["courant", "mixed"])
def test_tet_refusals(ctx, what):
    from dnlPython import _lib
    if what == "courant":
        m = samples.taylor_tet(n=2, nz=4, t_end=1e-6, dt_method=dnl.Explicit.Courant, context=ctx)
    else:
        m = samples.taylor_tet(n=2, nz=4, t_end=1e-6, precision="mixed", context=ctx)
    with pytest.raises(dnl.DnlgError) as e:
        m.solve()
    assert e.value.status == _lib.E_INVALID
