"""Failure statuses on the GPU (every case is forced from the host; every
kernel exits normally).

Each case stops the job with the oracle's status at the oracle's increment
and raises its Python exception; the job keeps the state of its last
completed cycle and stays failed.  In a batch the failed jobs stop while the
others run to their end time with results bit-identical to their solo runs
(Model.C:764-770, Element.C:560-581: the reference exits the process).
"""
import numpy as np
import pytest

from failure_models import CASES

pytestmark = pytest.mark.gpu


def _exc(case):
    import dnlPython as dnl
    return {"inverted": dnl._lib.NegativeJacobianError, "nan_speed": dnl._lib.NonFiniteError,
            "nr_limit": dnl._lib.NewtonRaphsonError, "crush": dnl._lib.NegativeJacobianError}[case]


@pytest.mark.parametrize("case", sorted(CASES))
def test_failure_status_fp64_matches_oracle(ctx, case):
    from oracle.bridge import OracleJob
    m = CASES[case](context=ctx)
    o = OracleJob(m.descriptors())
    st = o.solve(m._solver.endTime)
    print(f"{case}: oracle status {st} at increment {o.scalars()[2]}", flush=True)
    with pytest.raises(_exc(case)):
        m.solve()
    t, dt, inc = m.getScalars()
    to, dto, inco = o.scalars()
    # the engine forms the Jacobian of the end of cycle n (Explicit.C:245-246)
    # inside the element pass of cycle n + 1: an element that turns inside
    # out during the run is reported by that cycle, at the time of cycle n
    lag = 1 if case == "crush" else 0
    assert inc == inco + lag and t == pytest.approx(to, rel=1e-12, abs=0.0)
    with pytest.raises(_exc(case)):   # stays failed
        m.solve()
    m.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_failure_status_mixed(ctx, case):
    m = CASES[case](precision="mixed", context=ctx)
    with pytest.raises(_exc(case)):
        m.solve()
    print(f"{case} mixed: stopped at increment {m.getScalars()[2]}", flush=True)
    m.close()


def _healthy(ctx):
    from dnlPython import samples
    return [samples.taylor_axi(nr=6, nz=30, t_end=8e-6, precision="fp64", context=ctx),
            samples.taylor_3d(n=4, nz=12, t_end=8e-6, precision="mixed", track_tensors=False, context=ctx)]


def _state(m):
    import dnlPython as dnl
    return (m.getNodalDisplacements(), m.getIntPointField(dnl.Field.Stress),
            m.getIntPointField(dnl.Field.plasticStrain), m.getEnergies(), m.getScalars())


def test_failed_jobs_leave_the_batch_unchanged(ctx):
    import dnlPython as dnl
    solo = _healthy(ctx)
    for m in solo:
        m.solve()
    ref = [_state(m) for m in solo]
    failing = [CASES[c](context=ctx) for c in sorted(CASES)]
    mix = failing[:2] + _healthy(ctx) + failing[2:]
    statuses = dnl.batch_solve(mix, raise_on_error=False)
    names = sorted(CASES)
    for k, c in enumerate(names):
        j = k if k < 2 else k + 2
        assert statuses[j] == {"inverted": dnl._lib.E_NEG_JACOBIAN, "nan_speed": dnl._lib.E_NONFINITE,
                               "nr_limit": dnl._lib.E_NR_DIVERGED, "crush": dnl._lib.E_NEG_JACOBIAN}[c], c
    for m, r in zip(mix[2:4], ref):
        s = _state(m)
        assert statuses[mix.index(m)] == 0
        for a, b in zip(s, r):
            assert np.array_equal(np.asarray(a), np.asarray(b))
    for m in solo + mix:
        m.close()
