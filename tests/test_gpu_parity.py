"""FP64 engine vs the FP64 CPU oracle on the same models (GPU).

Tolerances (FP64 vs FP64): the engine contracts multiply-adds into FMAs and
uses the device libm, so results differ from the oracle by rounding that the
time loop accumulates.  Stated bounds: nodal displacement 1e-9 normwise
relative, IP fields 1e-7 normwise relative, time and increment count exact.
"""
import numpy as np
import pytest

from helpers import disp_err, rel_norm_err

pytestmark = pytest.mark.gpu

DISP_TOL = 1e-9
FIELD_TOL = 1e-7


def run_both(model, t_end):
    from oracle.bridge import OracleJob
    a = model.descriptors()
    o = OracleJob(a)
    assert o.solve(t_end) == 0
    model.solve(t_end)
    return o


def compare(model, o, fields=("plasticStrain", "T", "Stress", "density", "internalEnergy")):
    import dnlPython as dnl
    t, dt, inc = model.getScalars()
    to, dto, inco = o.scalars()
    assert inc == inco
    assert t == pytest.approx(to, rel=1e-12)
    _, d_ref, v_ref = o.nodes()
    err = disp_err(model.getNodalDisplacements(), d_ref)
    print(f"increments={inc} disp_err={err:.3e}")
    assert err <= DISP_TOL
    ids = {"plasticStrain": 0, "T": 5, "Stress": 11, "density": 6, "internalEnergy": 8}
    for f in fields:
        g = model.getIntPointField(getattr(dnl.Field, f))
        r = o.ip(ids[f])
        e = rel_norm_err(g, r)
        print(f"  {f}: {e:.3e}")
        assert e <= FIELD_TOL, f
    np.testing.assert_allclose(model.getEnergies(), o.energies(), rtol=1e-7, atol=1e-12)


def test_taylor_axi_short(ctx):
    from dnlPython import samples
    m = samples.taylor_axi(t_end=10e-6, context=ctx)
    o = run_both(m, 10e-6)
    compare(m, o)


def test_taylor_axi_full(ctx):
    from dnlPython import samples
    m = samples.taylor_axi(t_end=80e-6, context=ctx)
    o = run_both(m, 80e-6)
    compare(m, o)


def test_taylor_3d_small(ctx):
    from dnlPython import samples
    m = samples.taylor_3d(n=4, nz=10, t_end=20e-6, context=ctx)
    o = run_both(m, 20e-6)
    compare(m, o)


def test_necking_small(ctx):
    from dnlPython import samples
    m = samples.necking_3d(n=4, nz=12, t_end=20e-6, pull=5.0, context=ctx)
    o = run_both(m, 20e-6)
    compare(m, o)


# ---- mixed precision: FP32 strain / stress / forces, FP64 kinematics ------
# Bounds stated for the mixed paths against the FP64 oracle (normwise
# relative, max_i |a_i - ref_i| / max_i |ref_i|), both mixed modes:
#   nodal displacement 1e-5; equivalent plastic strain, temperature and
#   gammaCumulate 5e-5; energy balance (KE + IE) 5e-6;
#   stress and yield stress 1e-2: an integration point whose trial stress
#   sits within FP32 rounding of its yield stress can take the elastic branch
#   in one run and a tiny plastic step in the other, and then keeps a yield
#   stress from another cycle's strain rate; such flips dominate these two
#   fields (measured 2e-3 to 5e-3 on the 80 us Taylor impacts) while the
#   plastic strain, the temperature, the energies and the displacements stay
#   1e-7 to 1e-5.  gamma (the last increment) is reported, not bounded.
MIXED_DISP_TOL = 1e-5
_TOLS = dict(plasticStrain=5e-5, T=5e-5, gammaCumulate=5e-5, energy=5e-6, Stress=1e-2, yieldStress=1e-2)
MIXED_TOLS = {"mixed": _TOLS, "mixed-rr32": _TOLS}


def compare_mixed(model, o, prec):
    import dnlPython as dnl
    _, d_ref, _ = o.nodes()
    err = disp_err(model.getNodalDisplacements(), d_ref)
    out = {"disp": err}
    ids = {"plasticStrain": 0, "gamma": 2, "gammaCumulate": 3, "yieldStress": 4, "T": 5, "Stress": 11}
    for f, k in ids.items():
        out[f] = rel_norm_err(model.getIntPointField(getattr(dnl.Field, f)), o.ip(k))
    e_g, e_o = model.getEnergies(), o.energies()
    out["energy"] = float(abs(e_g[3] - e_o[3]) / abs(e_o[3]))
    print(f"{prec} errors:", {k: f"{v:.2e}" for k, v in out.items()})
    assert out["disp"] <= MIXED_DISP_TOL
    for k, tol in MIXED_TOLS[prec].items():
        assert out[k] <= tol, (k, out[k], tol)
    return out


MIXED_MODES = ["mixed", "mixed-rr32"]


@pytest.mark.parametrize("prec", MIXED_MODES)
def test_mixed_taylor_axi(ctx, prec):
    from dnlPython import samples
    m = samples.taylor_axi(t_end=80e-6, precision=prec, track_tensors=False, context=ctx)
    o = run_both(m, 80e-6)
    compare_mixed(m, o, prec)


@pytest.mark.parametrize("prec", MIXED_MODES)
def test_mixed_taylor_3d(ctx, prec):
    from dnlPython import samples
    m = samples.taylor_3d(n=6, nz=30, t_end=40e-6, precision=prec, track_tensors=False, context=ctx)
    o = run_both(m, 40e-6)
    compare_mixed(m, o, prec)


@pytest.mark.parametrize("prec", MIXED_MODES)
def test_mixed_necking(ctx, prec):
    from dnlPython import samples
    m = samples.necking_3d(n=4, nz=12, t_end=40e-6, pull=5.0, precision=prec, track_tensors=False,
                           context=ctx)
    o = run_both(m, 40e-6)
    compare_mixed(m, o, prec)


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed-rr32"])
def test_taylor_axi_published_results_on_gpu(ctx, prec):
    """The reference's Taylor-Axi sample (Samples/Taylor/Taylor-Axi/Taylor.py,
    4340 steel, 5 x 50, 287 m/s, 80 us) through the engine: published final
    length 26.52 mm and impact-face diameter 11.15 mm
    (Documentation/Samples-Impacts.tex:106-107), centre element plastic
    strain 1.78 and temperature 582.34 C."""
    from dnlPython import samples
    steel = dict(young=206000.0, poisson=0.3, density=7.83e-09, heatCapacity=4.6e+08, taylorQuinney=0.9,
                 A=806.0, B=614.0, C=0.0089, n=0.168, m=1.1, depsp0=1.0, Tm=1540.0, T0=20.0)
    m = samples.taylor_axi(nr=5, nz=50, radius=3.2, length=32.4, speed=287000.0, t_end=80e-6, material=steel,
                           precision=prec, context=ctx)
    m.solve()
    import dnlPython as dnl
    c = m.getNodalCoordinates()
    L_f, D_f = c[1].max(), 2 * c[0][:6].max()
    ep = m.getIntPointField(dnl.Field.plasticStrain)[0].mean()
    T = m.getIntPointField(dnl.Field.T)[0].mean()
    print(f"{prec}: L_f={L_f:.3f} D_f={D_f:.3f} ep={ep:.3f} T={T:.1f}")
    assert L_f == pytest.approx(26.52, rel=1e-3)
    assert D_f == pytest.approx(11.15, rel=1e-3)
    assert ep == pytest.approx(1.78, rel=1e-2)
    assert T == pytest.approx(582.34, rel=2e-3)
