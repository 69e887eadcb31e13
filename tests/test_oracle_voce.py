"""Voce + linear hardening law of the necking workloads in the CPU oracle
(no GPU).

sigma_y(p) = A + Q (1 - exp(-b p)) + H p and its slope Q b exp(-b p) + H
have no counterpart in the reference (its laws are JohnsonCookLaw.C,
ElastoplasticLaw.C, TabularLaw.C ...); the law is pinned here by what fixes
it independently of the oracle's code:
  * values computed with mpmath at 50 digits (a different evaluation path),
    at p = 0 (exactly A), near 0, mid-range and saturated;
  * the limits sigma_y(0) = A, slope(0) = Q b + H, slope(inf) = H and
    sigma_y(p) - H p -> A + Q;
  * the slope against central differences of the yield stress, and against
    mpmath's derivative;
  * no dependence on the plastic strain rate or the temperature;
  * the radial return (Element.C:485-601) with this law: the converged gamma
    is the root of the return equation found by brentq;
  * the reference's TabularLaw (TabularLaw.C, DiscreteFunction.C:188-308)
    sampled densely from the same curve reproduces a short necking solve to
    the table's interpolation error.
A misread term (sign of the exponent, Q b against Q / b, a dropped H) fails
the first three groups; a law that the return used differently from its
stated formula fails the last two.
"""
import ctypes

import mpmath
import numpy as np
import pytest
from scipy.optimize import brentq

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import F_GAMMA, F_PLASTIC_STRAIN, F_STRESS, F_YIELD_STRESS, OracleJob, load

A, Q, b, H = 80e6, 300e6, 16.93, 129.24e6   # the necking bar's law (MPa in Pa)
LAW_VOCE = 2                                # DNLG_LAW_VOCE_LINEAR


def _material(E=206e9, nu=0.3, rho=7830.0, cp=460.0, eta=0.9, T0=20.0):
    from dnlPython._lib import MaterialDesc
    m = MaterialDesc()
    m.young, m.poisson, m.density, m.heat_capacity, m.taylor_quinney, m.T0 = E, nu, rho, cp, eta, T0
    m.law = LAW_VOCE
    for k, v in enumerate((A, Q, b, H)):
        m.p[k] = v
    return m


M = _material()


def yld(p, dep=0.0, T=20.0):
    return load().oracle_law_yield(ctypes.addressof(M), p, dep, T)


def dyld(p, dep=0.0, T=20.0, dt=1e-8):
    return load().oracle_law_dyield(ctypes.addressof(M), p, dep, T, dt)


def _mp_yield(p):
    p = mpmath.mpf(p)
    return mpmath.mpf(A) + mpmath.mpf(Q) * (1 - mpmath.exp(-mpmath.mpf(b) * p)) + mpmath.mpf(H) * p


@pytest.mark.parametrize("p", [0.0, 1e-9, 1e-6, 3.7e-3, 0.05, 0.5, 3.0])
def test_voce_yield_against_mpmath(p):
    mpmath.mp.dps = 50
    assert yld(p) == pytest.approx(float(_mp_yield(p)), rel=4e-16, abs=0.0)


def test_voce_limits():
    assert yld(0.0) == A                                       # exactly A at p = 0
    assert dyld(0.0) == pytest.approx(Q * b + H, rel=1e-15)    # initial slope
    assert dyld(40.0) == pytest.approx(H, rel=1e-15)           # saturated: the linear term only
    assert yld(40.0) - H * 40.0 == pytest.approx(A + Q, rel=1e-12)
    # hand value: 80 + 300 (1 - e^-8.465) + 64.62 MPa at p = 0.5
    assert yld(0.5) == pytest.approx(444.5567852e6, rel=1e-9)


@pytest.mark.parametrize("p", [1e-6, 1e-3, 0.02, 0.2, 1.0, 2.5])
def test_voce_slope_against_differences_and_mpmath(p):
    h = 1e-6 * max(p, 1e-3)
    fd = (yld(p + h) - yld(p - h)) / (2 * h)
    assert dyld(p) == pytest.approx(fd, rel=1e-7)
    mpmath.mp.dps = 50
    assert dyld(p) == pytest.approx(float(mpmath.diff(_mp_yield, p)), rel=1e-14)


def test_voce_ignores_rate_and_temperature():
    for p in (0.0, 0.01, 0.7):
        y = yld(p)
        for dep, T in ((1e4, 20.0), (0.0, 900.0), (1e7, 1500.0)):
            assert yld(p, dep, T) == y
            assert dyld(p, dep, T, 3e-8) == dyld(p)


def test_voce_radial_return_brute_force_root():
    """One IP: the oracle's return (Element.C:485-601) against brentq on
    f(g) = S_tr - sqrt(3/2) 2G g - sigma_y(ep0 + sqrt(2/3) g)."""
    from test_oracle_kernels import radial_return
    G2 = M.young / (1 + M.poisson)
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(60):
        ep0 = float(rng.choice([0.0, rng.uniform(1e-4, 1.5)]))
        sy0 = yld(ep0) if ep0 > 0 else 0.0   # 0: the first-yield initialisation (Element.C:478-483)
        state = np.zeros(14)
        state[6], state[8], state[9], state[11] = ep0, sy0, rng.uniform(1e-7, 1e-4), 20.0
        dE = rng.normal(scale=2e-3, size=6)
        out = radial_return(M, dE, 0.0, 3e-8, state)
        dev = dE.copy()
        tr = (dE[0] + dE[3] + dE[5]) / 3
        dev[0] -= tr; dev[3] -= tr; dev[5] -= tr
        s_tr = G2 * dev
        S_tr = np.sqrt(1.5) * np.sqrt(s_tr[0] ** 2 + s_tr[3] ** 2 + s_tr[5] ** 2
                                      + 2 * (s_tr[1] ** 2 + s_tr[2] ** 2 + s_tr[4] ** 2))
        y_start = sy0 if ep0 > 0 else yld(1e-8)
        if S_tr <= y_start:
            continue

        def f(g):
            return S_tr - np.sqrt(1.5) * G2 * g - yld(ep0 + np.sqrt(2 / 3) * g)

        g_ref = brentq(f, 0.0, (S_tr - A) / (np.sqrt(1.5) * G2) * 1.0000001, xtol=1e-18, rtol=1e-15)
        assert out[9] == pytest.approx(g_ref, rel=1e-7, abs=1e-12)
        assert out[6] == pytest.approx(ep0 + np.sqrt(2 / 3) * out[9], rel=1e-14)
        # stored yield stress: the law at the last Newton iterate, at most one
        # step (|dgamma| < 1e-8, Element.C:485-601) from the final gamma
        assert abs(out[8] - yld(out[6])) <= dyld(out[6]) * np.sqrt(2 / 3) * 1e-8
        checked += 1
    assert checked >= 30


def test_voce_table_reproduces_the_voce_solve():
    """A TabularLaw through 20 000 points of the Voce curve (dense where it
    bends) gives the Voce solve of a plastic necking bar to the table's
    interpolation error."""
    def run(law):
        m = samples.necking_axi(nr=4, nz=20, pull=40.0, t_end=6e-6)
        m._material.setHardeningLaw(law)
        m._invalidate()
        o = OracleJob(m.descriptors(), threads=4)
        assert o.solve(6e-6) == 0
        return o

    voce = dnl.VoceLinearLaw("voce")
    voce.setParameters(A, Q, b, H)
    xs = np.concatenate([[0.0], np.geomspace(1e-10, 4.0, 20000)])
    f = dnl.DiscreteFunction("voce-table")
    for x in xs:
        f.add(x, A + Q * (1 - np.exp(-b * x)) + H * x)
    tab = dnl.TabularLaw("tab")
    tab.setFunction(f)
    a, t = run(voce), run(tab)
    ea, et = a.ip(F_PLASTIC_STRAIN), t.ip(F_PLASTIC_STRAIN)
    assert ea.max() > 1e-3          # the bar is plastic
    assert np.abs(ea - et).max() <= 1e-4 * ea.max()
    for fld in (F_STRESS, F_YIELD_STRESS, F_GAMMA):
        va, vt = a.ip(fld), t.ip(fld)
        assert np.abs(va - vt).max() <= 1e-4 * np.abs(va).max(), fld
    _, da, _ = a.nodes()
    _, dt_, _ = t.nodes()
    assert np.abs(da - dt_).max() <= 1e-6 * np.abs(da).max()
