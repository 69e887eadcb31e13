"""Serving loop (GPU): jobs submitted at different times, admitted by memory
footprint and advanced together in short steps finish with the same results
as the same job solved alone.  Per-job task lists and reductions make the
results bit-identical, in FP64 and in mixed precision."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs():
    from dnlPython import samples
    return [
        lambda ctx, p: samples.taylor_axi(6, 30, speed=200.0, t_end=12e-6, precision=p, context=ctx),
        lambda ctx, p: samples.taylor_3d(4, 12, speed=250.0, t_end=9e-6, precision=p, context=ctx),
        lambda ctx, p: samples.necking_axi(5, 40, pull=4.0, t_end=10e-6, precision=p, context=ctx),
        lambda ctx, p: samples.necking_3d(4, 10, pull=4.0, t_end=8e-6, precision=p, context=ctx),
        lambda ctx, p: samples.taylor_axi(10, 50, speed=280.0, t_end=6e-6, precision=p, context=ctx),
    ]


def _same(a, b):
    import dnlPython as dnl
    assert a.getScalars() == b.getScalars()
    np.testing.assert_array_equal(a.getNodalDisplacements(), b.getNodalDisplacements())
    np.testing.assert_array_equal(a.getNodalSpeeds(), b.getNodalSpeeds())
    np.testing.assert_array_equal(a.getIntPointField(dnl.Field.plasticStrain),
                                  b.getIntPointField(dnl.Field.plasticStrain))
    np.testing.assert_array_equal(a.getIntPointField(dnl.Field.Stress), b.getIntPointField(dnl.Field.Stress))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_server_results_equal_alone(ctx, prec):
    import dnlPython as dnl
    specs = _specs()
    alone = [f(ctx, prec) for f in specs]
    for m in alone:
        m.solve()
    srv = dnl.Server(ctx)
    models = [f(ctx, prec) for f in specs]
    tickets = [srv.submit(models[0]), srv.submit(models[1])]
    srv.step(7)
    tickets.append(srv.submit(models[2]))
    srv.step(5)
    tickets += [srv.submit(models[3]), srv.submit(models[4])]
    out = {}
    for _ in range(100000):
        running, queued = srv.step(16)
        for t in tickets:
            if t not in out and srv.poll(t)[0] == dnl.Server.DONE:
                out[t] = srv.take(t)
        if len(out) == len(tickets):
            break
    assert len(out) == len(tickets)
    for t, a in zip(tickets, alone):
        m, status = out[t]
        assert status == 0
        _same(a, m)
    srv.close()


def _footprint(model):
    from dnlPython import _lib
    a = model.descriptors()
    return _lib.load().dnlg_model_footprint_bytes(ctypes.byref(a.mesh), ctypes.byref(a.material),
                                                  ctypes.byref(a.bcs), ctypes.byref(a.params))


def test_server_admission_by_footprint():
    """With a budget just above the largest job's footprint, the running jobs
    never hold more than the budget (the largest job runs alone, smaller
    ones may share), all finish and match the alone results; a job larger
    than the whole budget is refused at submission.  Its own context, with
    no schedule table: the default 16-cycle graphs make a 32-cycle step
    stop inside the largest job's 79 cycles, so the test sees it running
    (a tuned 80-cycle graph would run it to its end within one step)."""
    import dnlPython as dnl
    from dnlPython import samples
    ctx = dnl.Context()
    ctx.clear_schedules()
    specs = _specs()[:3]
    alone = [f(ctx, "fp64") for f in specs]
    for m in alone:
        m.solve()
    models = [f(ctx, "fp64") for f in specs]
    fps = [_footprint(m) for m in models]
    budget = max(fps) + 1
    srv = dnl.Server(ctx, hbm_budget=budget)
    tickets = [srv.submit(m) for m in models]
    out = {}
    largest_alone = False
    for _ in range(100000):
        srv.step(32)
        states = [srv.poll(t)[0] if t not in out else None for t in tickets]
        held = sum(f for f, s in zip(fps, states) if s in (dnl.Server.RUNNING, dnl.Server.DONE))
        assert held <= budget
        big = int(np.argmax(fps))
        if states[big] == dnl.Server.RUNNING:
            assert sum(s == dnl.Server.RUNNING for s in states) == 1
            largest_alone = True
        for t, s in zip(tickets, states):
            if s == dnl.Server.DONE:
                out[t] = srv.take(t)
        if len(out) == len(tickets):
            break
    assert largest_alone
    for t, a in zip(tickets, alone):
        m, status = out[t]
        assert status == 0
        _same(a, m)
    big = samples.taylor_3d(12, 40, t_end=1e-6, context=ctx)
    assert _footprint(big) > budget
    with pytest.raises(dnl.DnlgError):
        srv.submit(big)
    srv.close()


def test_server_table_laws(ctx, tmp_path):
    """Jobs whose law carries a table (tabular points, a packed network)
    keep the server's own copy of it: after submission the caller's
    descriptors are dropped, and the results equal the jobs solved alone."""
    import gc

    import dnlPython as dnl
    from dnlPython import samples
    from helpers import ann_law, ann_network

    def tabular():
        f = dnl.DiscreteFunction("t")
        for x in np.concatenate([[0.0], np.geomspace(1e-9, 5.0, 200)]):
            f.add(x, 90e6 + 292e6 * x ** 0.31)
        law = dnl.TabularLaw("tab")
        law.setFunction(f)
        return law

    def make(kind):
        m = samples.taylor_axi(5, 25, speed=220.0, t_end=8e-6, context=ctx)
        m._material.setHardeningLaw(tabular() if kind == "tab" else ann_law(tmp_path / "n.npz", ann_network(seed=5)))
        m._invalidate()
        return m

    alone = [make("tab"), make("ann")]
    for m in alone:
        m.solve()
    srv = dnl.Server(ctx)
    tickets = [srv.submit(make("tab")), srv.submit(make("ann"))]
    gc.collect()
    out = {}
    for _ in range(100000):
        srv.step(16)
        for t in tickets:
            if t not in out and srv.poll(t)[0] == dnl.Server.DONE:
                out[t] = srv.take(t)
        if len(out) == len(tickets):
            break
    assert len(out) == len(tickets)
    for t, a in zip(tickets, alone):
        m, status = out[t]
        assert status == 0
        _same(a, m)
    srv.close()
