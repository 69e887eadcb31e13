"""The "Direct" stress integration (Element::computeStressDirect,
Element.C:625-768) in the CPU oracle (no GPU).  With a linear hardening law
(power law with n = 1, no rate or temperature terms) the linearised return
is exact, so it must reproduce the Newton-Raphson return (to the latter's
1e-8 tolerance on gamma); with Johnson-Cook it is only a first-order
approximation and the two differ."""
import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import F_PLASTIC_STRAIN, F_STRESS, OracleJob


def _run(method, law=None, t_end=10e-6):
    m = samples.taylor_axi(nr=5, nz=25, t_end=t_end)
    if law is not None:
        m._material.setHardeningLaw(law)
        m._invalidate()
    m.setStressComputation(method)
    o = OracleJob(m.descriptors(), threads=4)
    assert o.solve(t_end) == 0
    return o


def test_direct_equals_newton_raphson_for_linear_hardening():
    def linear():
        law = dnl.ElastoplasticLaw("lin")
        law.setParameters(120e6, 800e6, 1.0)
        return law

    a, b = _run("NewtonRaphson", linear()), _run("Direct", linear())
    assert a.scalars()[2] == b.scalars()[2]
    ea, eb = a.ip(F_PLASTIC_STRAIN), b.ip(F_PLASTIC_STRAIN)
    assert ea.max() > 0.05
    assert np.abs(ea - eb).max() <= 1e-6 * ea.max()
    sa, sb = a.ip(F_STRESS), b.ip(F_STRESS)
    assert np.abs(sa - sb).max() <= 1e-6 * np.abs(sa).max()


def test_direct_differs_from_newton_raphson_for_johnson_cook():
    a, b = _run("NewtonRaphson"), _run("Direct")
    ea, eb = a.ip(F_PLASTIC_STRAIN), b.ip(F_PLASTIC_STRAIN)
    assert np.abs(ea - eb).max() > 1e-3 * ea.max()


def test_stress_method_names():
    m = samples.taylor_axi(nr=2, nz=4)
    with pytest.raises(dnl.DnlgError):
        m.setStressComputation("Bissection")
