"""Global energies by the device's compensated reduction (GPU).

Model::getTotalKineticEnergy (Model.C:612-625) sums m |v|^2 / 2 over the
nodes.  The engine sums every term with Neumaier's compensation in a fixed
order, so:
  * the kinetic energy equals math.fsum (the exactly rounded sum) of the
    per-node terms formed from the read-back masses and speeds, to a few ulp
    (the terms themselves can differ by an ulp: the device contracts
    v.v into fused multiply-adds);
  * two reads give the same bits, and a job's energies are the same bits
    alone and in a batch;
  * internal energy and plastic work equal the FP64 oracle's to the FP64
    parity bound of the trajectory.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ke_terms(model):
    v = model.getNodalSpeeds()
    m = model.getNodalMasses()
    return m * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / 2.0


def test_kinetic_energy_is_the_exactly_rounded_sum(ctx):
    from dnlPython import samples
    m = samples.taylor_3d(n=20, nz=100, t_end=2e-6, precision="fp64", track_tensors=False, context=ctx)
    m.initSolve()
    # t = 0: v = (0, -227, 0) on every node, terms formed exactly as on the host
    terms = _ke_terms(m)
    exact = math.fsum(terms)
    ke = m.getEnergies()[0]
    print(f"t=0: KE {ke!r} fsum {exact!r} naive {float(np.sum(terms))!r}")
    assert abs(ke - exact) <= 2.3e-16 * exact
    m.solve()
    terms = _ke_terms(m)
    exact = math.fsum(terms)
    ke = m.getEnergies()[0]
    print(f"t=2us: KE {ke!r} fsum {exact!r} rel {abs(ke - exact) / exact:.2e}")
    assert abs(ke - exact) <= 4 * 2.3e-16 * exact
    m.close()


def test_energies_reproducible_alone_and_in_batch(ctx):
    import dnlPython as dnl
    from dnlPython import samples

    def jobs():
        return [samples.taylor_axi(nr=6, nz=30, t_end=6e-6, precision="fp64", context=ctx),
                samples.taylor_3d(n=4, nz=12, t_end=6e-6, precision="mixed", track_tensors=False, context=ctx)]

    alone = jobs()
    for m in alone:
        m.solve()
    mix = jobs()
    dnl.batch_solve(mix)
    for a, b in zip(alone, mix):
        ea, eb = a.getEnergies(), b.getEnergies()
        assert np.array_equal(ea, a.getEnergies())
        assert np.array_equal(ea, eb), (ea, eb)
        assert ea[3] == ea[0] + ea[1] or abs(ea[3] - (ea[0] + ea[1])) <= 2.3e-16 * abs(ea[3])
    for m in alone + mix:
        m.close()


def test_internal_energy_and_plastic_work_against_oracle(ctx):
    from dnlPython import samples
    from oracle.bridge import OracleJob
    m = samples.taylor_3d(n=4, nz=10, t_end=10e-6, precision="fp64", context=ctx)
    o = OracleJob(m.descriptors(), threads=4)
    assert o.solve(10e-6) == 0
    m.solve()
    e, eo = m.getEnergies(), o.energies()
    print("engine", e, "oracle", eo)
    assert e[2] > 0.0
    np.testing.assert_allclose(e, eo, rtol=1e-9)
    m.close()
