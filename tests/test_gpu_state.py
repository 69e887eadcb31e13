"""Checkpoint / restart on the GPU, and parity with the oracle from a
developed plastic state at the full sizes of the 3D workloads.

  * restart: a job that takes another job's state at t1 continues exactly
    as that job continues (FP64 and mixed: the same bits); a mixed quad job
    clears the xz / yz stress rows of a state it takes (its kernel carries
    none);
  * the 120 000-hex Taylor impact (80 us in total) run by the engine to
    40 us, deep in its plastic phase, then continued by both the engine and
    the CPU oracle from the engine's state for about 45 cycles: the FP64
    engine within the FP64 parity bounds, the mixed engine within the mixed
    bounds (the window's displacement increment, normwise relative);
  * the 998 400-hex necking bar the same way for 8 cycles.
Tolerances (stated): FP64 window displacement 1e-9, IP fields 1e-7; mixed
window displacement 1e-5, plastic strain 5e-5.
"""
import numpy as np
import pytest

from helpers import disp_err, rel_norm_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_restart_continues_bit_for_bit(ctx, prec):
    import dnlPython as dnl
    from dnlPython import samples

    def make():
        return samples.taylor_3d(n=6, nz=30, t_end=16e-6, precision=prec, track_tensors=(prec == "fp64"),
                                 context=ctx)
    a = make()
    a.solve(7e-6)
    st = a.getState()
    b = make()
    b.setState(st)
    a.solve()
    b.solve()
    assert a.getScalars() == b.getScalars()
    assert np.array_equal(a.getNodalCoordinates(), b.getNodalCoordinates())
    for f in (dnl.Field.Stress, dnl.Field.plasticStrain, dnl.Field.T, dnl.Field.gammaCumulate):
        assert np.array_equal(a.getIntPointField(f), b.getIntPointField(f)), f
    assert np.array_equal(a.getEnergies(), b.getEnergies())
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["mixed", "mixed-rr32"])
def test_restart_quad_mixed_out_of_plane_stress(ctx, prec):
    """The mixed kernel of the plane and axisymmetric quads carries no xz /
    yz stress: a state handed to such a job has those rows cleared, and the
    job continues with the same bits as the original continuing."""
    import dnlPython as dnl
    from dnlPython import samples

    def make():
        return samples.taylor_axi(8, 40, t_end=16e-6, precision=prec, track_tensors=False, context=ctx)
    a = make()
    a.solve(7e-6)
    st = a.getState()
    assert not np.any(st["stress"][[2, 4]])
    st["stress"][[2, 4]] = 1.0e3
    b = make()
    b.setState(st)
    assert not np.any(b.getState()["stress"][[2, 4]])
    a.solve()
    b.solve()
    assert a.getScalars() == b.getScalars()
    assert np.array_equal(a.getNodalCoordinates(), b.getNodalCoordinates())
    for f in (dnl.Field.Stress, dnl.Field.plasticStrain, dnl.Field.T):
        assert np.array_equal(a.getIntPointField(f), b.getIntPointField(f)), f
    a.close()
    b.close()


def _window_parity(model_fp64, t1, t2, prec_models, threads=16):
    """FP64 engine to t1; its state to the oracle and to a model of each
    precision; all advance to t2; returns the errors per precision."""
    import dnlPython as dnl
    from oracle.bridge import OracleJob
    model_fp64.solve(t1)
    st = model_fp64.getState()
    d1 = model_fp64.getNodalDisplacements()
    o = OracleJob(model_fp64.descriptors(), threads=threads)
    assert o.set_state(st) == 0
    assert o.solve(t2) == 0
    _, d_ref, _ = o.nodes()
    out = {}
    for prec, m in prec_models.items():
        if m is not model_fp64:
            m.setState(st)
        m.solve(t2)
        if prec == "fp64":
            assert m.getScalars()[2] == o.scalars()[2]
        inc = m.getNodalDisplacements() - d1
        e = {"window_disp": disp_err(inc, d_ref - d1), "disp": disp_err(m.getNodalDisplacements(), d_ref)}
        for name, k in (("plasticStrain", 0), ("T", 5), ("Stress", 11)):
            e[name] = rel_norm_err(m.getIntPointField(getattr(dnl.Field, name)), o.ip(k))
        print(f"{prec}: cycles {o.scalars()[2] - st['increment']} from t={t1 * 1e6:g} us, "
              + ", ".join(f"{k} {v:.2e}" for k, v in e.items()), flush=True)
        out[prec] = e
    return out


def test_taylor_3d_120k_midrun_against_oracle(ctx):
    from dnlPython import samples
    kw = dict(n=20, nz=100, t_end=80e-6, track_tensors=False, context=ctx)
    m64 = samples.taylor_3d(precision="fp64", **kw)
    mmx = samples.taylor_3d(precision="mixed", **kw)
    e = _window_parity(m64, 40e-6, 40.2e-6, {"fp64": m64, "mixed": mmx})
    assert e["fp64"]["window_disp"] <= 1e-9 and e["fp64"]["disp"] <= 1e-9
    for k in ("plasticStrain", "T", "Stress"):
        assert e["fp64"][k] <= 1e-7, k
    assert e["mixed"]["window_disp"] <= 1e-5 and e["mixed"]["plasticStrain"] <= 5e-5
    m64.close()
    mmx.close()


def test_necking_3d_1m_midrun_against_oracle(ctx):
    from dnlPython import samples
    kw = dict(n=40, nz=208, pull=1.0, t_end=50e-6, track_tensors=False, context=ctx)
    m64 = samples.necking_3d(precision="fp64", **kw)
    assert m64.getElementsNumber() == 998400
    mmx = samples.necking_3d(precision="mixed", **kw)
    e = _window_parity(m64, 20e-6, 20.25e-6, {"fp64": m64, "mixed": mmx})
    assert e["fp64"]["window_disp"] <= 1e-9 and e["fp64"]["disp"] <= 1e-9
    for k in ("T", "Stress"):
        assert e["fp64"][k] <= 1e-7, k
    assert e["mixed"]["window_disp"] <= 1e-5
    m64.close()
    mmx.close()
