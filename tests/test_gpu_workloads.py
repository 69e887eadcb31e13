"""The measured workloads as GPU tests: the headline impact at its full size
and length, seeded subsets of the batched sweep and of the serving stream.

  * Taylor impact 3D, 120 000 hex8, the full 80 us in each mixed mode and in
    FP64: mixed vs FP64 nodal displacement <= 1e-5 (normwise relative),
    plastic strain <= 1e-4 of its maximum, the energy balance (KE + IE
    against 1/2 M v0^2, the rigid wall does no work) within 2e-5 of KE0 of
    the FP64 run's and below 1 % in every precision.
  * Batched sweep: 12 seeded jobs of the sweep's generator (axisymmetric and
    3D impacts, 2k-20k elements, Johnson-Cook constants and speed drawn),
    advanced as one batch to 4 us, 4.3 us and 8 us.  Every job equals its
    solo run with the same stops bit for bit; over the 4 -> 4.3 us window,
    in the plastic phase, each job's batch result matches the oracle
    continued from the job's own state: FP64 window displacement <= 1e-9
    and IP fields <= 1e-7, mixed window displacement <= 1e-5 and plastic
    strain <= 5e-5.
  * Sweep jobs run to 80 us, mixed vs FP64 against the FP64 solution's own
    sensitivity: the same FP64 job with the impact speed times (1 +- 2^-24),
    one FP32 rounding.  Hex jobs: mixed within 1e-5.  Axisymmetric jobs of
    the sweep are ill-conditioned -- a one-ulp change of the speed moves
    the FP64 answer by more than 2e-6 -- and their mixed difference stays
    within 10x that movement.
  * Serving stream: 8 seeded jobs of the serving mix (axisymmetric / 3D x
    impact / necking, 1k-20k elements, end times 5-15 us) submitted at
    Poisson-drawn server steps while the server advances the others: each
    job equals its solo run bit for bit; the latencies' p50 / p99 are
    printed.
"""
import time

import numpy as np
import pytest

from helpers import disp_err, rel_norm_err

pytestmark = pytest.mark.gpu


def _ke0(m, speed=227.0):
    return 0.5 * float(m.getNodalMasses().sum()) * speed ** 2


def test_taylor_3d_120k_full_run_mixed_vs_fp64(ctx):
    import dnlPython as dnl
    from dnlPython import samples
    res = {}
    for prec in ("fp64", "mixed", "mixed-rr32"):
        m = samples.taylor_3d(n=20, nz=100, t_end=80e-6, precision=prec, track_tensors=False, context=ctx)
        m.solve()
        e = m.getEnergies()
        res[prec] = dict(disp=m.getNodalDisplacements(), ep=m.getIntPointField(dnl.Field.plasticStrain),
                         balance=(e[0] + e[1] - _ke0(m)) / _ke0(m), cycles=m.getIncrement())
        m.close()
    ref = res["fp64"]
    print(f"fp64: {ref['cycles']} cycles, energy balance {ref['balance']:.3e}, ep max {ref['ep'].max():.3f}",
          flush=True)
    assert abs(ref["balance"]) < 1e-2
    for prec in ("mixed", "mixed-rr32"):
        r = res[prec]
        d = disp_err(r["disp"], ref["disp"])
        ep = float(np.abs(r["ep"] - ref["ep"]).max() / ref["ep"].max())
        db = abs(r["balance"] - ref["balance"])
        print(f"{prec}: {r['cycles']} cycles, disp {d:.2e}, plastic strain {ep:.2e}, "
              f"energy balance {r['balance']:.3e} (vs FP64 {db:.2e})", flush=True)
        assert d <= 1e-5
        assert ep <= 1e-4
        assert db <= 2e-5 and abs(r["balance"]) < 1e-2


def _arrays(m):
    import dnlPython as dnl
    return [m.getNodalDisplacements(), m.getNodalSpeeds(), m.getIntPointField(dnl.Field.Stress),
            m.getIntPointField(dnl.Field.plasticStrain), m.getIntPointField(dnl.Field.T),
            np.array(m.getScalars()), m.getEnergies()]


def _same(a, b):
    for x, y in zip(_arrays(a), _arrays(b)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_sweep_subset_batch_equals_solo_and_oracle(ctx, prec):
    import dnlPython as dnl
    from dnlPython import samples
    from oracle.bridge import OracleJob
    jobs = samples.sweep_params(12, seed=7, lo=2000, hi=20000)
    assert {j["kind"] for j in jobs} == {"axi", "3d"}
    t1, t2, t_end = 4e-6, 4.3e-6, 8e-6

    def make(j):
        return samples.build_sweep_job(j, t_end=t_end, precision=prec, context=ctx)
    batch = [make(j) for j in jobs]
    print(f"{prec}: {len(jobs)} jobs, {sum(m.getElementsNumber() for m in batch)} elements", flush=True)
    assert dnl.batch_solve(batch, upTo=t1) == [0] * len(jobs)
    states = [m.getState() for m in batch]
    d1 = [m.getNodalDisplacements() for m in batch]
    assert dnl.batch_solve(batch, upTo=t2) == [0] * len(jobs)
    at_t2 = [_arrays(m) for m in batch]
    worst = {}
    for j, m, st, d in zip(jobs, batch, states, d1):
        o = OracleJob(m.descriptors(), threads=16)
        assert o.set_state(st) == 0
        assert o.solve(t2) == 0
        assert o.scalars()[2] == m.getIncrement()
        _, d_ref, _ = o.nodes()
        e = {"window_disp": disp_err(m.getNodalDisplacements() - d, d_ref - d)}
        for name, k in (("plasticStrain", 0), ("T", 5), ("Stress", 11)):
            e[name] = rel_norm_err(m.getIntPointField(getattr(dnl.Field, name)), o.ip(k))
        for k, v in e.items():
            worst[k] = max(worst.get(k, 0.0), v)
        if prec == "fp64":
            assert e["window_disp"] <= 1e-9, (j, e)
            assert max(e["plasticStrain"], e["T"], e["Stress"]) <= 1e-7, (j, e)
        else:
            assert e["window_disp"] <= 1e-5, (j, e)
            assert e["plasticStrain"] <= 5e-5, (j, e)
    print(f"{prec}: worst over the jobs vs the oracle, 4 -> 4.3 us: "
          + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()), flush=True)
    assert dnl.batch_solve(batch) == [0] * len(jobs)
    for j, m, a2 in zip(jobs, batch, at_t2):
        s = make(j)
        s.solve(t1)
        s.solve(t2)
        for x, y in zip(_arrays(s), a2):
            assert np.array_equal(x, y)
        s.solve()
        _same(s, m)
        s.close()
        m.close()


@pytest.mark.parametrize("idx", [5, 8, 16, 31])
def test_sweep_job_full_run_mixed_vs_fp64_sensitivity(ctx, idx):
    from dnlPython import samples
    job = samples.sweep_params(64, seed=20261203)[idx]

    def run(prec, scale=1.0):
        j = dict(job, speed=job["speed"] * scale)
        m = samples.build_sweep_job(j, t_end=80e-6, precision=prec, context=ctx)
        m.solve()
        d = m.getNodalDisplacements().copy()
        m.close()
        return d

    ref = run("fp64")
    sens = max(disp_err(run("fp64", 1.0 + s * 2.0 ** -24), ref) for s in (1, -1))
    err = disp_err(run("mixed"), ref)
    print(f"job {idx} {job['kind']} {job['shape']}: FP64 one-ulp sensitivity {sens:.2e}, mixed vs FP64 {err:.2e}",
          flush=True)
    if job["kind"] == "3d":
        assert err <= 1e-5
    else:
        assert sens > 2e-6
        assert err <= 10.0 * sens


def test_serving_stream_jobs_equal_solo(ctx):
    import dnlPython as dnl
    from dnlPython import samples
    specs = samples.serve_params(8, seed=12, lo=1000, hi=20000)
    rng = np.random.default_rng(13)
    for s in specs:
        s["t_end"] = float(rng.uniform(5e-6, 15e-6))
    assert len({(s["kind"], s["load"]) for s in specs}) == 4
    arrival = np.floor(np.cumsum(rng.exponential(3.0, size=len(specs)))).astype(int)
    srv = dnl.Server(ctx)
    models = [samples.build_serve_job(s, context=ctx) for s in specs]
    tickets, submitted, out = {}, {}, {}
    step = 0
    while len(out) < len(specs):
        for i in np.nonzero(arrival == step)[0]:
            tickets[srv.submit(models[i])] = i
            submitted[i] = (step, time.perf_counter())
        srv.step(64)
        step += 1
        for t, i in tickets.items():
            if i not in out and srv.poll(t)[0] == dnl.Server.DONE:
                m, status = srv.take(t)
                assert status == 0
                out[i] = (m, step - submitted[i][0], time.perf_counter() - submitted[i][1])
        assert step < 100000
    srv.close()
    lat_steps = np.array([out[i][1] for i in range(len(specs))])
    lat_s = np.array([out[i][2] for i in range(len(specs))])
    print(f"serving stream: {len(specs)} jobs, {sum(m.getElementsNumber() for m in models)} elements, "
          f"latency p50 {np.percentile(lat_s, 50):.3f} s / {np.percentile(lat_steps, 50):.0f} steps, "
          f"p99 {np.percentile(lat_s, 99):.3f} s / {np.percentile(lat_steps, 99):.0f} steps", flush=True)
    for i, s in enumerate(specs):
        solo = samples.build_serve_job(s, context=ctx)
        solo.solve()
        _same(solo, out[i][0])
        solo.close()
        out[i][0].close()
