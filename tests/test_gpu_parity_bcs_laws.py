"""Engine vs oracle on the paths the cycle control treats specially (GPU):
ramped prescribed speeds (the field0 BC pass runs every cycle), constant
speeds across solve segments (the step is clipped at each segment end, so
the speed BC values change with it), and the power hardening law.

Tolerances as tests/test_gpu_parity.py: FP64 displacement 1e-9, IP fields
1e-7; mixed displacement 1e-5 (normwise relative)."""
import numpy as np
import pytest

from helpers import disp_err
from test_gpu_parity import compare, run_both

pytestmark = pytest.mark.gpu


def _ramp(model, ramp, speed_scale, t0, t1):
    import dnlPython as dnl
    from dnlPython import _lib
    n = 0
    for b, _members in model._constant:
        if b.kind == _lib.BC_SPEED:
            rf = dnl.RampFunction("ramp")
            rf.setFunction(ramp, t0, t1)
            b.setValue(0.0, speed_scale, 0.0)
            b.setFunction(rf)
            n += 1
    assert n == 1
    model._invalidate()
    return model


@pytest.mark.parametrize("ramp", [1, 2, 5])  # Linear, Sinus, Soft (RampFunction.h:31-39)
def test_ramped_pull_axi_fp64(ctx, ramp):
    import dnlPython as dnl
    from dnlPython import samples
    # the ramp density integrates to 1 over [0, 20 us]: speed_scale is the
    # imposed end displacement (20 um over 20 us, ~1 m/s average)
    m = _ramp(samples.necking_axi(5, 40, t_end=12e-6, context=ctx), ramp, 20e-6, 0.0, 20e-6)
    o = run_both(m, 12e-6)
    compare(m, o)


def test_ramped_pull_3d_mixed(ctx):
    from dnlPython import samples
    m = _ramp(samples.necking_3d(4, 10, t_end=10e-6, precision="mixed", context=ctx), 5, 30e-6, 0.0, 15e-6)
    o = run_both(m, 10e-6)
    _, d_ref, _ = o.nodes()
    err = disp_err(m.getNodalDisplacements(), d_ref)
    print(f"mixed ramped pull disp_err={err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_speed_bc_across_segments(ctx, prec):
    """Three solve calls on the engine and the oracle; each call clips the
    step at its end, which changes u = v dt on the pulled nodes."""
    from dnlPython import samples
    from oracle.bridge import OracleJob
    m = samples.necking_axi(5, 40, pull=4.0, t_end=12e-6, precision=prec, context=ctx)
    o = OracleJob(m.descriptors())
    for t in (3e-6, 7.5e-6, 12e-6):
        assert o.solve(t) == 0
        m.solve(t)
    if prec == "fp64":
        compare(m, o)
    else:
        _, d_ref, _ = o.nodes()
        assert disp_err(m.getNodalDisplacements(), d_ref) <= 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_power_law_taylor_axi(ctx, prec):
    """ElastoplasticLaw sigma_y = A + B ep^n (ElastoplasticLaw.C:110-121)."""
    import dnlPython as dnl
    from dnlPython import samples
    m = samples.taylor_axi(6, 30, speed=200.0, t_end=10e-6, precision=prec, context=ctx)
    law = dnl.ElastoplasticLaw("power")
    law.setParameters(90e6, 292e6, 0.31)
    mat = m._material
    mat.setHardeningLaw(law)
    m._invalidate()
    o = run_both(m, 10e-6)
    if prec == "fp64":
        compare(m, o)
    else:
        _, d_ref, _ = o.nodes()
        err = disp_err(m.getNodalDisplacements(), d_ref)
        print(f"mixed power-law disp_err={err:.3e}")
        assert err <= 1e-5
    assert np.max(m.getIntPointField(dnl.Field.plasticStrain)) > 0.01


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_direct_stress_integration(ctx, kind):
    """setStressComputation("Direct") (Element.C:625-768) against the oracle."""
    from dnlPython import samples
    if kind == "axi":
        m = samples.taylor_axi(nr=5, nz=25, t_end=10e-6, context=ctx)
    else:
        m = samples.taylor_3d(n=4, nz=12, t_end=10e-6, context=ctx)
    m.setStressComputation("Direct")
    o = run_both(m, 10e-6)
    compare(m, o)


def test_direct_stress_integration_has_no_fp32_form(ctx):
    """mixed-rr32 (the radial return in FP32) refuses the Direct integration."""
    import dnlPython as dnl
    from dnlPython import samples
    m = samples.taylor_axi(nr=3, nz=8, precision="mixed-rr32", context=ctx)
    m.setStressComputation("Direct")
    with pytest.raises(dnl.DnlgError):
        m.solve()


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_direct_stress_integration_mixed(ctx, kind):
    """Direct on the mixed path (FP32 strain / stress, the linearised return
    in FP64) against the FP64 oracle, at the mixed tolerances."""
    from dnlPython import samples
    from test_gpu_parity import compare_mixed
    if kind == "axi":
        m = samples.taylor_axi(nr=5, nz=25, t_end=10e-6, precision="mixed", track_tensors=False, context=ctx)
    else:
        m = samples.taylor_3d(n=4, nz=12, t_end=10e-6, precision="mixed", track_tensors=False, context=ctx)
    m.setStressComputation("Direct")
    o = run_both(m, 10e-6)
    compare_mixed(m, o, "mixed")


def _tabular(points):
    import dnlPython as dnl
    f = dnl.DiscreteFunction("table")
    for x, y in points:
        f.add(x, y)
    law = dnl.TabularLaw("tab")
    law.setFunction(f)
    return law


def test_tabular_law(ctx):
    """TabularLaw (TabularLaw.C) through a table of the copper impact's
    Johnson-Cook curve at room temperature, against the oracle."""
    from dnlPython import samples
    xs = np.concatenate([[0.0], np.geomspace(1e-9, 5.0, 300)])
    law = _tabular(zip(xs, 90e6 + 292e6 * xs ** 0.31))
    m = samples.taylor_axi(nr=5, nz=25, t_end=10e-6, context=ctx)
    m._material.setHardeningLaw(law)
    m._invalidate()
    o = run_both(m, 10e-6)
    compare(m, o)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_tabular_law_errors(ctx, prec):
    """Out of the table: the job stops with DNLG_E_TABLE_RANGE (the reference
    exits, DiscreteFunction.C:195-196), in both precisions that run the
    law; mixed-rr32 (an FP32 return) is refused."""
    import dnlPython as dnl
    from dnlPython import _lib, samples
    m = samples.taylor_axi(nr=4, nz=20, t_end=8e-6, precision=prec, context=ctx)
    m._material.setHardeningLaw(_tabular([(0.0, 120e6), (0.01, 125e6)]))
    m._invalidate()
    with pytest.raises(dnl.DnlgError) as e:
        m.solve()
    assert e.value.status == _lib.E_TABLE_RANGE
    m = samples.taylor_axi(nr=4, nz=20, precision="mixed-rr32", context=ctx)
    m._material.setHardeningLaw(_tabular([(0.0, 120e6), (1.0, 125e6)]))
    m._invalidate()
    with pytest.raises(dnl.DnlgError) as e:
        m.solve()
    assert e.value.status == _lib.E_INVALID


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_tabular_law_mixed(ctx, kind):
    """The tabular law on the mixed path (its FP64 return from the
    reference's starting state) against the FP64 oracle, at the mixed
    tolerances: a dense table of the Johnson-Cook strain hardening."""
    import dnlPython as dnl
    from dnlPython import samples
    from test_gpu_parity import compare_mixed
    pts = [(e, 90e6 + 292e6 * e ** 0.31) for e in np.concatenate([[0.0], np.geomspace(1e-8, 4.0, 400)])]
    if kind == "axi":
        m = samples.taylor_axi(nr=5, nz=25, t_end=10e-6, precision="mixed", track_tensors=False, context=ctx)
    else:
        m = samples.taylor_3d(n=4, nz=12, t_end=10e-6, precision="mixed", track_tensors=False, context=ctx)
    m._material.setHardeningLaw(_tabular(pts))
    m._invalidate()
    o = run_both(m, 10e-6)
    compare_mixed(m, o, "mixed")
    assert np.max(m.getIntPointField(dnl.Field.plasticStrain)) > 0.05


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_ann_law(ctx, tmp_path, kind):
    """ANN2SigLaw (ANN2SigLaw.C:173-245): a rate- and temperature-dependent
    network of 7 and 5 neurons, against the oracle."""
    import dnlPython as dnl
    from dnlPython import samples
    from helpers import ann_law, ann_network
    if kind == "axi":
        m = samples.taylor_axi(nr=5, nz=25, t_end=10e-6, context=ctx)
    else:
        m = samples.taylor_3d(n=4, nz=12, t_end=10e-6, context=ctx)
    m._material.setHardeningLaw(ann_law(tmp_path / "net.npz", ann_network(seed=5)))
    m._invalidate()
    o = run_both(m, 10e-6)
    compare(m, o)
    assert np.max(m.getIntPointField(dnl.Field.plasticStrain)) > 0.05


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_ann_law_mixed(ctx, tmp_path, kind):
    """The network law on the mixed path (its FP64 return) against the FP64
    oracle at the mixed tolerances, alone and in a batch with a
    Johnson-Cook mixed job that stays bit-identical to its solo run."""
    import dnlPython as dnl
    from dnlPython import samples
    from helpers import ann_law, ann_network
    from test_gpu_parity import compare_mixed
    from oracle.bridge import OracleJob
    mk = samples.taylor_axi if kind == "axi" else samples.taylor_3d
    kw = dict(nr=5, nz=25) if kind == "axi" else dict(n=4, nz=12)
    m = mk(t_end=10e-6, precision="mixed", track_tensors=False, context=ctx, **kw)
    m._material.setHardeningLaw(ann_law(tmp_path / "net.npz", ann_network(seed=5)))
    m._invalidate()
    jc = mk(t_end=10e-6, precision="mixed", track_tensors=False, context=ctx, **kw)
    solo = mk(t_end=10e-6, precision="mixed", track_tensors=False, context=ctx, **kw)
    o = OracleJob(m.descriptors())
    assert o.solve(10e-6) == 0
    dnl.batch_solve([m, jc])
    solo.solve()
    compare_mixed(m, o, "mixed")
    assert np.max(m.getIntPointField(dnl.Field.plasticStrain)) > 0.05
    np.testing.assert_array_equal(jc.getNodalDisplacements(), solo.getNodalDisplacements())


def test_ann_law_in_a_batch_with_another_law(ctx, tmp_path):
    """One batch, one network-law job (its own element kernel variant) and
    one Johnson-Cook job of the same element type: each matches its oracle
    run, the Johnson-Cook job bit for bit its own solo run."""
    import dnlPython as dnl
    from dnlPython import samples
    from helpers import ann_law, ann_network
    from oracle.bridge import OracleJob
    a = samples.taylor_axi(nr=5, nz=25, t_end=10e-6, context=ctx)
    a._material.setHardeningLaw(ann_law(tmp_path / "net.npz", ann_network(seed=9, h1=15, h2=7)))
    a._invalidate()
    b = samples.taylor_axi(nr=4, nz=30, t_end=10e-6, context=ctx)
    solo = samples.taylor_axi(nr=4, nz=30, t_end=10e-6, context=ctx)
    oa, ob = OracleJob(a.descriptors()), OracleJob(b.descriptors())
    assert oa.solve(10e-6) == 0 and ob.solve(10e-6) == 0
    dnl.batch_solve([a, b])
    solo.solve()
    compare(a, oa)
    compare(b, ob)
    np.testing.assert_array_equal(b.getNodalDisplacements(), solo.getNodalDisplacements())


def test_ann_law_errors(ctx, tmp_path):
    """Job creation refuses mixed-rr32 (an FP32 return) for a network law,
    and packed networks whose sizes, table length or scaling ranges are
    invalid."""
    import dnlPython as dnl
    from dnlPython import _lib, samples
    from helpers import ann_law, ann_network

    def attempt(precision="fp64", edit=None):
        m = samples.taylor_axi(nr=3, nz=8, precision=precision, context=ctx)
        law = ann_law(tmp_path / "net.npz", ann_network())
        if edit is not None:
            law._table = edit(law._table.copy())
        m._material.setHardeningLaw(law)
        m._invalidate()
        with pytest.raises(dnl.DnlgError) as e:
            m.solve()
        assert e.value.status == _lib.E_INVALID

    attempt("mixed-rr32")

    def too_wide(t):
        t[0] = 65
        return t

    def short(t):
        return t[:-1]

    def empty_range(t):
        t[-1] = t[-5]  # maxEntries[3] = minEntries[3]
        return t

    def bad_base(t):
        t[-9] = 0.0  # logBase
        return t

    for edit in (too_wide, short, empty_range, bad_base):
        attempt(edit=edit)


def test_ann_fortran_law(ctx):
    """ANNFortranLaw (the fixed 3-7-4-1 network, yield stress in MPa; the
    SI copper bar at 0.05 m/s is then plastic from the first cycles)
    against the oracle: its output bias is negative, which the packed
    network's validation must accept."""
    import dnlPython as dnl
    from dnlPython import samples
    m = samples.taylor_axi(nr=5, nz=25, speed=0.05, t_end=2e-6, context=ctx)
    m._material.setHardeningLaw(dnl.ANNFortranLaw("f"))
    m._invalidate()
    o = run_both(m, 2e-6)
    compare(m, o)
    assert np.max(m.getIntPointField(dnl.Field.plasticStrain)) > 1e-5
