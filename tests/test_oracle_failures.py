"""Failure statuses of the CPU oracle (no GPU): where the reference calls
exit(-1) (Errors.C:157,196), the restatement returns a status and keeps the
state of the last completed cycle.

Pinned against the reference's behaviour:
  * a negative Jacobian in the pre-loop's reference pass stops before any
    increment (Model.C:754-775 runs before the loop, Explicit.C:143-150);
  * a NaN initial speed propagates into the coordinates in the first cycle
    and is caught at its end (the engine's check; the reference has none and
    would run on with NaN);
  * the Newton limit stops at the first return that exceeds it;
  * an inside-out element during the run stops at that cycle's Jacobian.
"""
import pytest

from dnlPython import _lib
from failure_models import CASES
from oracle.bridge import OracleJob

EXPECTED = {"inverted": (_lib.E_NEG_JACOBIAN, 0), "nan_speed": (_lib.E_NONFINITE, 1),
            "nr_limit": (_lib.E_NR_DIVERGED, 2), "crush": (_lib.E_NEG_JACOBIAN, 9)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_failure_status(case):
    m = CASES[case]()
    o = OracleJob(m.descriptors())
    st = o.solve(m._solver.endTime)
    t, dt, inc = o.scalars()
    assert (st, inc) == EXPECTED[case]
    assert o.status() == st
    assert o.solve(m._solver.endTime) == st     # a failed job stays failed


def test_oracle_default_newton_limit_is_the_reference_250():
    """nr_max_iter = 0 keeps the reference's 250: the Taylor impact that the
    limit 1 stops runs through."""
    m = CASES["nr_limit"]()
    m._solver.setMaxNewtonIterations(0)
    o = OracleJob(m.descriptors())
    assert o.solve(m._solver.endTime) == 0
