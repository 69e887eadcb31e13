"""Power-iteration time step (Model.C:942-1009) on the GPU against the FP64
oracle (GPU).  Same bounds as the Courant parity tests: nodal
displacement 1e-9 normwise relative, IP fields 1e-7, time and increment
count exact; mixed precision 1e-5."""
import os

import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from test_gpu_parity import MIXED_DISP_TOL, compare, compare_mixed, run_both

pytestmark = pytest.mark.gpu

PI = dnl.Explicit.PowerIteration


@pytest.mark.parametrize("freq", [10, 1])
def test_pi_taylor_axi(ctx, freq):
    m = samples.taylor_axi(t_end=10e-6, freq=freq, dt_method=PI, context=ctx)
    o = run_both(m, 10e-6)
    compare(m, o)


def test_pi_taylor_3d(ctx):
    m = samples.taylor_3d(n=4, nz=12, t_end=12e-6, dt_method=PI, context=ctx)
    o = run_both(m, 12e-6)
    compare(m, o)


def test_pi_necking_axi(ctx):
    """Samples/BarNecking/BarNecking-Axi selects the power iteration
    (BarNecking.py:132): a pulled bar with speed BCs."""
    m = samples.necking_axi(nr=6, nz=40, t_end=10e-6, pull=4.0, dt_method=PI, context=ctx)
    o = run_both(m, 10e-6)
    compare(m, o)


def test_pi_mixed(ctx):
    m = samples.taylor_axi(t_end=40e-6, freq=1, dt_method=PI, precision="mixed", track_tensors=False, context=ctx)
    o = run_both(m, 40e-6)
    e = compare_mixed(m, o, "mixed")
    assert e["disp"] <= MIXED_DISP_TOL


def test_pi_iterations_finished_by_the_host_give_the_same_results():
    """With no iteration slot inside the launch sequences every time-step
    cycle stalls and the host runs the whole iteration: results are
    bit-identical to the default (two slots, host only when they do not
    suffice)."""
    def run(slots):
        old = os.environ.get("DNLG_PI_SLOTS")
        os.environ["DNLG_PI_SLOTS"] = str(slots)
        try:
            c = dnl.Context(0)
        finally:
            if old is None:
                del os.environ["DNLG_PI_SLOTS"]
            else:
                os.environ["DNLG_PI_SLOTS"] = old
        m = samples.taylor_axi(nr=6, nz=30, t_end=15e-6, dt_method=PI, context=c)
        m.solve()
        return m
    a, b = run(2), run(0)
    assert a.getScalars() == b.getScalars()
    np.testing.assert_array_equal(a.getNodalDisplacements(), b.getNodalDisplacements())
    np.testing.assert_array_equal(a.getIntPointField(dnl.Field.Stress), b.getIntPointField(dnl.Field.Stress))


def test_pi_and_courant_jobs_in_one_batch(ctx):
    """A batch mixing both time-step methods gives each job its result alone."""
    def models():
        return [samples.taylor_axi(nr=6, nz=30, t_end=12e-6, dt_method=PI, context=ctx),
                samples.taylor_axi(nr=6, nz=30, t_end=12e-6, context=ctx),
                samples.taylor_3d(n=4, nz=12, t_end=10e-6, dt_method=PI, context=ctx)]
    alone = models()
    for m in alone:
        m.solve()
    batch = models()
    dnl.batch_solve(batch)
    for a, b in zip(alone, batch):
        assert a.getScalars() == b.getScalars()
        np.testing.assert_array_equal(a.getNodalDisplacements(), b.getNodalDisplacements())


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_taylor_axi_sample_with_its_power_iteration_step_on_gpu(ctx, prec):
    """Samples/Taylor/Taylor-Axi/Taylor.py as published (power-iteration
    step every increment): final length 26.52 mm, impact-face diameter
    11.15 mm (Documentation/Samples-Impacts.tex:106-107)."""
    steel = dict(young=206000.0, poisson=0.3, density=7.83e-09, heatCapacity=4.6e+08, taylorQuinney=0.9,
                 A=806.0, B=614.0, C=0.0089, n=0.168, m=1.1, depsp0=1.0, Tm=1540.0, T0=20.0)
    m = samples.taylor_axi(nr=5, nz=50, radius=3.2, length=32.4, speed=287000.0, t_end=80e-6, material=steel,
                           freq=1, dt_method=PI, precision=prec, context=ctx)
    m.solve()
    c = m.getNodalCoordinates()
    L_f, D_f = c[1].max(), 2 * c[0][:6].max()
    print(f"{prec}: L_f={L_f:.3f} mm D_f={D_f:.3f} mm")
    assert L_f == pytest.approx(26.52, rel=1e-3)
    assert D_f == pytest.approx(11.15, rel=1e-3)


PIU = dnl.Explicit.PowerIterationUnder


@pytest.mark.parametrize("kind", ["axi", "3d", "plane"])
def test_pi_under(ctx, kind):
    """PowerIterationUnder (Solver.C:212-216: the power iteration on the
    one-point under-integrated element stiffness) against the oracle, the
    step recomputed every increment, FP64."""
    if kind == "axi":
        m = samples.taylor_axi(t_end=10e-6, freq=1, dt_method=PIU, context=ctx)
    elif kind == "3d":
        m = samples.taylor_3d(n=4, nz=12, t_end=10e-6, freq=1, dt_method=PIU, context=ctx)
    else:
        m = samples.necking_plane(nx=8, ny=24, t_end=10e-6, pull=4.0, freq=1, dt_method=PIU, context=ctx)
    o = run_both(m, 10e-6)
    compare(m, o)


def test_pi_under_mixed_and_larger_step(ctx):
    """The under-integrated stiffness leaves out the hourglass modes, so its
    highest frequency is lower and the step larger than the fully
    integrated one's on the same mesh; the mixed path against the oracle."""
    m = samples.taylor_axi(t_end=20e-6, freq=1, dt_method=PIU, precision="mixed", track_tensors=False,
                           context=ctx)
    full = samples.taylor_axi(t_end=20e-6, freq=1, dt_method=PI, precision="mixed", track_tensors=False,
                              context=ctx)
    o = run_both(m, 20e-6)
    compare_mixed(m, o, "mixed")
    full.solve()
    assert m.getScalars()[2] < full.getScalars()[2]
