"""Launch schedules and the autotuner on the GPU.

A schedule changes how the element kernels and the corrector are laid out
(threads per block, waves per SIMD, slot batching, cycles per graph), never
what is computed: the same model gives the same bits under every schedule
the kernels are built for, in FP64 and both mixed paths, for hexahedra and
axisymmetric quads.  The autotuner measures schedules by advancing the job,
so a tuned job ends with the same bits as an untuned one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BLOCKS = {"hex": (64, 128, 256), "axi": (128, 256, 512)}
ALT = {"fp64": 2, "mixed": 5, "mixed-rr32": 5}


def _model(ctx, family, prec, t_end=6e-6):
    from dnlPython import samples
    if family == "hex":
        return samples.taylor_3d(n=4, nz=12, t_end=t_end, precision=prec, track_tensors=False, context=ctx)
    return samples.taylor_axi(nr=8, nz=40, t_end=t_end, precision=prec, track_tensors=False, context=ctx)


def _state(m):
    import dnlPython as dnl
    return [m.getNodalDisplacements(), m.getNodalSpeeds(), m.getIntPointField(dnl.Field.Stress),
            m.getIntPointField(dnl.Field.plasticStrain), m.getIntPointField(dnl.Field.T),
            np.array(m.getScalars())]


@pytest.mark.parametrize("family", ["hex", "axi"])
@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed-rr32"])
def test_every_schedule_gives_the_same_bits(ctx, family, prec):
    ref = _model(ctx, family, prec)
    ref.solve()
    r = _state(ref)
    default = ref.getSchedule()
    scheds = [dict(elem_block=b, waves=w) for b in BLOCKS[family] for w in (0, ALT[prec])]
    scheds += [dict(corr_batch=c) for c in ((2, 4, 8) if prec != "fp64" else (2, 4))]
    scheds += [dict(graph_cycles=g) for g in (20, 80)]
    for sch in scheds:
        m = _model(ctx, family, prec)
        m.setSchedule(**sch)
        got = m.getSchedule()
        assert all(got[k] == v or (k == "waves" and v == 0) for k, v in sch.items()), (sch, got)
        m.solve()
        for a, b in zip(_state(m), r):
            assert np.array_equal(a, b), (sch, default)
        m.close()
    ref.close()


def test_autotune_leaves_results_unchanged(ctx):
    """Both models stop at 5 us (the step before it is clipped to land
    there) and then run to the end; the tuned one runs its trials in
    between, each an advance toward the end time."""
    untuned = _model(ctx, "axi", "mixed", t_end=40e-6)
    untuned.solve(5e-6)
    untuned.solve()
    m = _model(ctx, "axi", "mixed", t_end=40e-6)
    m.solve(5e-6)
    sched, ms, ms0 = m.autotune(cycles_per_trial=40, rounds=1)
    print("tuned", sched, f"{ms * 1e3:.1f} us/cycle (starting schedule {ms0 * 1e3:.1f})", flush=True)
    assert m.getSchedule() == sched and ms > 0 and ms0 > 0
    m.solve()
    for a, b in zip(_state(m), _state(untuned)):
        assert np.array_equal(a, b)
    m.close()
    untuned.close()
