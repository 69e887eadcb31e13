"""Checkpoint / restart of the CPU oracle (no GPU): a job that takes the
state of another at t1 continues exactly (bit for bit) as that job
continues, for the axisymmetric and the 3D impact, with and without the
strain tensors.  The state record is therefore complete: nothing that the
next cycles read is left out."""
import numpy as np
import pytest

from dnlPython import samples
from oracle.bridge import F_GAMMA_CUMULATE, F_PLASTIC_STRAIN, F_STRESS, F_TEMPERATURE, OracleJob


def _builders():
    return {
        "axi": lambda track: samples.taylor_axi(nr=5, nz=25, t_end=12e-6, track_tensors=track),
        "3d": lambda track: samples.taylor_3d(n=4, nz=10, t_end=12e-6, track_tensors=track),
    }


@pytest.mark.parametrize("kind", ["axi", "3d"])
@pytest.mark.parametrize("track", [False, True])
def test_oracle_restart_continues_bit_for_bit(kind, track):
    make = _builders()[kind]
    a = OracleJob(make(track).descriptors(), threads=4)
    assert a.solve(5e-6) == 0
    st = a.get_state()
    assert st["time"] == pytest.approx(5e-6, rel=1e-15) and st["increment"] > 0
    b = OracleJob(make(track).descriptors(), threads=4)
    assert b.set_state(st) == 0
    assert a.solve(12e-6) == 0 and b.solve(12e-6) == 0
    assert a.scalars() == b.scalars()
    ca, da, va = a.nodes()
    cb, db, vb = b.nodes()
    assert np.array_equal(ca, cb) and np.array_equal(va, vb)
    # the displacement accumulator restarts as coords - initial coords
    assert np.abs(da - db).max() <= 1e-12 * np.abs(da).max()
    for f in (F_PLASTIC_STRAIN, F_GAMMA_CUMULATE, F_TEMPERATURE, F_STRESS):
        assert np.array_equal(a.ip(f), b.ip(f)), f
    assert np.array_equal(a.energies(), b.energies())
    assert a.ip(F_PLASTIC_STRAIN).max() > 0.01   # through the plastic phase


def test_oracle_state_round_trip():
    a = OracleJob(samples.taylor_axi(nr=4, nz=20, t_end=4e-6).descriptors())
    assert a.solve(3e-6) == 0
    s1 = a.get_state()
    b = OracleJob(samples.taylor_axi(nr=4, nz=20, t_end=4e-6).descriptors())
    assert b.set_state(s1) == 0
    s2 = b.get_state()
    for k, v in s1.items():
        assert np.array_equal(np.asarray(v), np.asarray(s2[k])), k
