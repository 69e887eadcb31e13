"""Batched solve: uneven jobs advanced together give the same results as
each job alone (GPU).  FP64 jobs use per-job fixed-order reductions, so the
results are expected bit-identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _models(ctx, prec):
    from dnlPython import samples
    return [samples.taylor_axi(6, 30, speed=200.0, t_end=15e-6, precision=prec, context=ctx),
            samples.taylor_3d(4, 12, speed=250.0, t_end=12e-6, precision=prec, context=ctx),
            samples.taylor_axi(10, 50, speed=280.0, t_end=10e-6, precision=prec, context=ctx),
            samples.necking_3d(4, 10, t_end=9e-6, pull=4.0, precision=prec, context=ctx)]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_equals_alone(ctx, prec):
    import dnlPython as dnl
    alone = _models(ctx, prec)
    for m in alone:
        m.solve()
    mixed = _models(ctx, prec)
    dnl.batch_solve(mixed)
    for a, b in zip(alone, mixed):
        assert a.getScalars() == b.getScalars()
        np.testing.assert_array_equal(a.getNodalDisplacements(), b.getNodalDisplacements())
        np.testing.assert_array_equal(a.getIntPointField(dnl.Field.plasticStrain),
                                      b.getIntPointField(dnl.Field.plasticStrain))
        np.testing.assert_array_equal(a.getIntPointField(dnl.Field.Stress), b.getIntPointField(dnl.Field.Stress))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_branches_on_one_stream_give_the_same_bits(prec):
    """The 2D and 3D jobs of a batch run as two branches of each cycle on
    two streams; on one stream (DNLG_BRANCHES=0) the results are the same
    bits, with and without captured graphs."""
    import os
    import dnlPython as dnl

    def run(branches, graph_cycles):
        env = {"DNLG_BRANCHES": branches, "DNLG_GRAPH_CYCLES": graph_cycles}
        old = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            c = dnl.Context(0)
        finally:
            for k, v in old.items():
                if v is None:
                    del os.environ[k]
                else:
                    os.environ[k] = v
        ms = _models(c, prec)
        dnl.batch_solve(ms)
        return [(m.getScalars(), m.getNodalDisplacements(), m.getIntPointField(dnl.Field.Stress)) for m in ms]

    ref = run("0", "16")
    for branches, graph_cycles in (("1", "16"), ("1", "0")):
        for (sa, da, fa), (sb, db, fb) in zip(ref, run(branches, graph_cycles)):
            assert sa == sb
            np.testing.assert_array_equal(da, db)
            np.testing.assert_array_equal(fa, fb)


def test_batch_in_segments(ctx):
    """Advancing a batch in several solve calls equals one call."""
    import dnlPython as dnl
    one = _models(ctx, "fp64")
    dnl.batch_solve(one)
    seg = _models(ctx, "fp64")
    for t in (3e-6, 7e-6, None):
        dnl.batch_solve(seg, upTo=t)
    for a, b in zip(one, seg):
        ta, _, _ = a.getScalars()
        tb, _, _ = b.getScalars()
        assert ta == pytest.approx(tb, rel=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_advance_in_cycle_chunks(ctx, prec):
    """dnlg_batch_advance: advancing by short cycle caps until every job
    reaches its end time gives bit-identical results to one batch_solve
    (each stop closes the last cycle; the next call resumes it)."""
    import dnlPython as dnl
    one = _models(ctx, prec)
    dnl.batch_solve(one)
    chunked = _models(ctx, prec)
    for _ in range(100000):
        dnl.batch_solve(chunked, max_cycles=37)
        if all(m.getCurrentTime() >= m._solver.endTime for m in chunked):
            break
    for a, b in zip(one, chunked):
        assert a.getScalars() == b.getScalars()
        np.testing.assert_array_equal(a.getNodalDisplacements(), b.getNodalDisplacements())
        np.testing.assert_array_equal(a.getIntPointField(dnl.Field.Stress), b.getIntPointField(dnl.Field.Stress))
