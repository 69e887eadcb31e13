"""ANN2SigLaw (ANN2SigLaw.C:137-245) in the CPU oracle (no GPU): the yield
stress against a forward pass on the weights as the .npz file stores them
(catches a transposed or misread matrix, a dropped bias, a wrong input or
output scaling, the rate cut-off), the slope against central differences
of the yield stress, and a constant network's solve against the same solve
with a constant power law."""
import ctypes

import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from helpers import ann_forward, ann_law, ann_network
from oracle.bridge import (F_PLASTIC_STRAIN, F_PLASTIC_STRAIN_RATE, F_STRESS, F_TEMPERATURE, F_YIELD_STRESS,
                           OracleJob, load)

ETA, RHO, CP = 0.9, 8960.0, 383.0


def _material(law):
    mat = dnl.Material("m")
    mat.taylorQuinney, mat.density, mat.heatCapacity = ETA, RHO, CP
    mat.setHardeningLaw(law)
    return mat.descriptor()


@pytest.mark.parametrize("h1,h2", [(7, 5), (3, 11), (1, 1)])
def test_ann_yield_matches_the_file_layout_forward_pass(tmp_path, h1, h2):
    net = ann_network(seed=h1 * 100 + h2, h1=h1, h2=h2)
    d = _material(ann_law(tmp_path / "net.npz", net))
    lib = load()
    rng = np.random.default_rng(1)
    for _ in range(200):
        ep, T = rng.uniform(0.0, 2.0), rng.uniform(20.0, 1083.0)
        rate = float(np.exp(rng.uniform(-3.0, 14.0)))  # below and above logBase = 1
        got = lib.oracle_law_yield(ctypes.byref(d), ep, rate, T)
        assert got == pytest.approx(ann_forward(net, ep, rate, T), rel=1e-13)


def test_ann_slope_is_the_yield_stress_derivative(tmp_path):
    """hard = dY/dep + dY/d(dep) / dtime + eta / (rho Cp) Y dY/dT: each part
    by central differences of the oracle's yield stress; two time steps
    separate the rate term.  At or below logBase the rate input is 0 and
    the reference keeps the rate term with depsp = logBase: the slope there
    equals the slope just above logBase (minEntries[1] = 0)."""
    net = ann_network(seed=3)
    d = _material(ann_law(tmp_path / "net.npz", net))
    lib = load()
    Y = lambda e, r, T: lib.oracle_law_yield(ctypes.byref(d), e, r, T)  # noqa: E731
    H = lambda e, r, T, dt: lib.oracle_law_dyield(ctypes.byref(d), e, r, T, dt)  # noqa: E731
    for ep, rate, T in ((0.05, 2e3, 150.0), (0.4, 5e4, 400.0), (1.2, 30.0, 800.0)):
        he, hr, hT = 1e-6, rate * 1e-6, 1e-3
        dYe = (Y(ep + he, rate, T) - Y(ep - he, rate, T)) / (2 * he)
        dYr = (Y(ep, rate + hr, T) - Y(ep, rate - hr, T)) / (2 * hr)
        dYT = (Y(ep, rate, T + hT) - Y(ep, rate, T - hT)) / (2 * hT)
        for dt in (1e-8, 3e-7):
            want = dYe + dYr / dt + ETA / (RHO * CP) * Y(ep, rate, T) * dYT
            assert H(ep, rate, T, dt) == pytest.approx(want, rel=1e-6)
        # the rate part alone, from the difference of two time steps
        diff = H(ep, rate, T, 1e-8) - H(ep, rate, T, 3e-7)
        assert diff == pytest.approx(dYr * (1 / 1e-8 - 1 / 3e-7), rel=1e-6)
    # below the rate cut-off
    for dt in (1e-8, 3e-7):
        below = H(0.3, 0.25, 300.0, dt)
        above = H(0.3, 1.0 + 1e-12, 300.0, dt)
        assert below == pytest.approx(above, rel=1e-9)
        assert Y(0.3, 0.25, 300.0) == Y(0.3, 1.0, 300.0)


def test_ann_file_errors(tmp_path):
    net = ann_network()
    bad = dict(net)
    del bad["maxEntries"]
    with pytest.raises(dnl.DnlgError):
        ann_law(tmp_path / "a.npz", bad)
    bad = dict(net)
    bad["w2"] = net["w2"].ravel()
    with pytest.raises(dnl.DnlgError):
        ann_law(tmp_path / "b.npz", bad)
    bad = dict(net)
    bad["w2"] = net["w2"][:-1]
    with pytest.raises(dnl.DnlgError):
        ann_law(tmp_path / "c.npz", bad)
    law = dnl.ANN2SigLaw("x")
    assert law.getNumberOfParameters() == 9 and law.getParameterName(8) == "maxEntries"
    with pytest.raises(dnl.DnlgError):
        _material(law)  # no parameters read


def test_constant_network_solve_is_the_perfectly_plastic_solve(tmp_path):
    """w3 = 0: the network's yield stress is the constant
    (max3 - min3) b3 + min3 and its slope 0, so the solve is the one with
    the power law A = that constant, B = 0, bit for bit."""
    c = 0.4
    net = ann_network(constant=c)
    y0 = (net["maxEntries"][3] - net["minEntries"][3]) * c + net["minEntries"][3]

    def run(law):
        m = samples.taylor_axi(nr=4, nz=20, t_end=8e-6)
        m._material.setHardeningLaw(law)
        m._invalidate()
        o = OracleJob(m.descriptors(), threads=4)
        assert o.solve(8e-6) == 0
        return o

    power = dnl.ElastoplasticLaw("p")
    power.setParameters(y0, 0.0, 1.0)
    a, b = run(power), run(ann_law(tmp_path / "c.npz", net))
    assert a.ip(F_PLASTIC_STRAIN).max() > 0.01
    np.testing.assert_array_equal(a.ip(F_PLASTIC_STRAIN), b.ip(F_PLASTIC_STRAIN))
    np.testing.assert_array_equal(a.ip(F_STRESS), b.ip(F_STRESS))
    np.testing.assert_array_equal(a.nodes()[1], b.nodes()[1])


def test_hardening_network_solve(tmp_path):
    """A rate- and temperature-dependent network on the copper impact: the
    solve completes, plastic strain and temperature rise, and at the points
    that flowed in the last cycle the stored yield stress is the network's
    at the stored state (to 1e-3: the stored temperature is the one after
    the cycle's heat update)."""
    m = samples.taylor_axi(nr=4, nz=20, t_end=8e-6)
    m._material.setHardeningLaw(ann_law(tmp_path / "n.npz", ann_network(seed=5)))
    m._invalidate()
    net = ann_network(seed=5)
    o = OracleJob(m.descriptors(), threads=4)
    assert o.solve(8e-6) == 0
    ep, dep = o.ip(F_PLASTIC_STRAIN).ravel(), o.ip(F_PLASTIC_STRAIN_RATE).ravel()
    T, sy = o.ip(F_TEMPERATURE).ravel(), o.ip(F_YIELD_STRESS).ravel()
    assert ep.max() > 0.05 and T.max() > 20.0 + 1.0
    live = np.flatnonzero(dep > 0)
    assert len(live) > 10
    want = np.array([ann_forward(net, ep[g], dep[g], T[g]) for g in live])
    np.testing.assert_allclose(sy[live], want, rtol=1e-3)


def _fortran_form(law, ep, rate, T, dtime):
    """ANNFortranLaw's own expressions (ANNFortranLaw.C:76-114, 180-266) on
    the class's constants, which keep the reference's signs: layer 1 as
    1 + exp(A x + a), layer 2 as exp(C / zb + c), and the slope's rounded
    constants 90.348959964501 = scale / 10.8198 and 2.036574406340 =
    scale / 480."""
    x = np.array([ep, np.log(rate) / law._rate_div if rate > 1 else 0.0, (T - law._T0) / law._T_div])
    if rate <= 1:
        rate = 1.0
    A, a, C, c, E = (np.array(v) for v in (law._A, law._a, law._C, law._c, law._E))
    za = np.exp(A @ x + a)
    zb = 1 + za
    zc = np.exp(C @ (1 / zb) + c)
    y = E @ (1 / (1 + zc)) + law._e
    zd = E * (zc / (1 + zc) ** 2)
    ze = za / zb ** 2
    zf = (-(C.T @ zd)) * ze
    yd = -(A.T @ zf)
    yld = law._scale * y + law._offset
    hard = (law._scale * yd[0] + 90.348959964501 * yd[1] / rate / dtime
            + ETA / (RHO * CP) * yld * 2.036574406340 * yd[2])
    return yld, hard


def test_fortran_network_is_the_ann2sig_law_with_its_weights():
    """ANNFortranLaw packed as an ANN2SigLaw network gives the reference's
    closed-form yield stress and slope (the sign flips of the packing, the
    input scalings and the output scale)."""
    law = dnl.ANNFortranLaw("f")
    law.setParameters("ignored")
    d = _material(law)
    lib = load()
    rng = np.random.default_rng(7)
    for _ in range(200):
        ep, T = rng.uniform(0.0, 1.0), rng.uniform(20.0, 500.0)
        rate = float(np.exp(rng.uniform(-2.0, 10.0)))
        dtime = float(10 ** rng.uniform(-9, -6))
        y, h = _fortran_form(law, ep, rate, T, dtime)
        assert lib.oracle_law_yield(ctypes.byref(d), ep, rate, T) == pytest.approx(y, rel=1e-13)
        assert lib.oracle_law_dyield(ctypes.byref(d), ep, rate, T, dtime) == pytest.approx(h, rel=1e-9, abs=1e-9)


def test_elastic_law_is_the_zero_yield_return():
    """ElasticLaw's yield stress and slope are 0 (ElasticLaw.C:67-78) and
    the return treats it like any law (Element.C:470-480): after the solve
    every integration point's stress is a pure pressure, to the deviatoric
    stress 2G x 1e-8 that the return's tolerance on gamma leaves."""
    m = samples.taylor_axi(nr=4, nz=20, t_end=2e-6)
    m._material.setHardeningLaw(dnl.ElasticLaw("e"))
    m._invalidate()
    d = m._material.descriptor()
    lib = load()
    assert lib.oracle_law_yield(ctypes.byref(d), 0.3, 1e3, 300.0) == 0.0
    assert lib.oracle_law_dyield(ctypes.byref(d), 0.3, 1e3, 300.0, 1e-8) == 0.0
    o = OracleJob(m.descriptors(), threads=4)
    assert o.solve(2e-6) == 0
    sig = o.ip(F_STRESS)  # [xx, xy, xz, yy, yz, zz]
    p = (sig[0] + sig[3] + sig[5]) / 3
    scale = np.abs(p).max()
    assert scale > 0
    two_g = m._material.getShearModulus() * 2
    for c, iso in ((0, True), (1, False), (2, False), (3, True), (4, False), (5, True)):
        np.testing.assert_allclose(sig[c], p if iso else 0.0, rtol=0, atol=2 * two_g * 1e-8)
