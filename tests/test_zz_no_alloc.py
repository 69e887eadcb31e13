"""dnlg_job_energies_static, the entry point that allocates nothing (its
reduction partials live in a fixed device scratch array), against
dnlg_job_energies on the same jobs (GPU, through the C ABI): the same bits,
on two calls in a row with the second on a smaller job, so that nothing of
the first call's scratch contents leaks into the second."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _energies(ctx, name, model):
    out = np.full(4, np.nan)
    st = getattr(ctx.lib, name)(model._job, out.ctypes.data_as(_dp))
    return st, out


def test_static_scratch_energies_equal_the_allocating_entry_point(ctx):
    from dnlPython import samples
    big = samples.taylor_3d(n=20, nz=100, t_end=1e-6, precision="mixed", track_tensors=False, context=ctx)
    small = samples.taylor_axi(nr=6, nz=30, t_end=3e-6, precision="fp64", context=ctx)
    big.solve()
    small.solve()
    print(f"jobs: {big.getNodesNumber()} nodes / {big.getElementsNumber()} elements, then "
          f"{small.getNodesNumber()} / {small.getElementsNumber()}", flush=True)
    results = []
    for m in (big, small):
        st_s, e_s = _energies(ctx, "dnlg_job_energies_static", m)
        st_a, e_a = _energies(ctx, "dnlg_job_energies", m)
        assert st_s == 0 and st_a == 0
        print("static", e_s, "allocating", e_a, flush=True)
        assert np.array_equal(e_s, e_a)
        assert e_s[0] > 0.0 and np.all(np.isfinite(e_s))
        results.append(e_s)
    assert not np.array_equal(results[0], results[1])
    # and again in the other order: the small job right after the large one
    st, e = _energies(ctx, "dnlg_job_energies_static", big)
    assert st == 0 and np.array_equal(e, results[0])
    st, e = _energies(ctx, "dnlg_job_energies_static", small)
    assert st == 0 and np.array_equal(e, results[1])
    big.close()
    small.close()
