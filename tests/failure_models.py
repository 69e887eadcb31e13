"""Models that stop with each of the engine's failure statuses, forced from
the host so that every kernel still exits normally (shared by the oracle
and GPU failure tests).

  inverted    one hex built with its bottom and top faces swapped: negative
              detJ in the pre-loop's reference Jacobian (Model.C:764-770,
              detected at Element3D.C:98-102)
  nan_speed   an initial BoundarySpeed with a NaN component: the nodal state
              is non-finite after the first cycle
  nr_limit    the radial return's Newton limit set to 1 iteration (the
              reference's is 250, Element.C:560-581): the first return that
              needs a second Newton step stops the job
  crush       a 3000 m/s Taylor impact: impact-face quads turn inside out
              during the run (negative detJ at a cycle)
"""
from dnlPython import samples


def inverted(precision="fp64", context=None):
    m = samples.taylor_3d(n=2, nz=4, t_end=4e-6, precision=precision, context=context)
    c = m._elem_conn[0]
    c[5] = c[5][[4, 5, 6, 7, 0, 1, 2, 3]]
    m._invalidate()
    return m


def nan_speed(precision="fp64", context=None):
    m = samples.taylor_axi(nr=4, nz=20, t_end=4e-6, precision=precision, context=context)
    bc, _ = m._initial[0]
    bc.setValue(0, float("nan"), 0)
    m._invalidate()
    return m


def nr_limit(precision="fp64", context=None):
    m = samples.taylor_axi(nr=4, nz=20, t_end=4e-6, precision=precision, context=context)
    m._solver.setMaxNewtonIterations(1)
    return m


def crush(precision="fp64", context=None):
    return samples.taylor_axi(nr=4, nz=20, speed=3000.0, t_end=20e-6, precision=precision, context=context)


CASES = {"inverted": inverted, "nan_speed": nan_speed, "nr_limit": nr_limit, "crush": crush}
