#!/usr/bin/env python3
"""Benchmark of the explicit-solve engine on one MI355X.

Default workload (--workload taylor3d): 3D Taylor impact of a full copper
cylinder, O-grid cross-section of 20 x 20 + 4 x 20 x 10 cells x 100 layers =
120 000 8-node hexes (125 341 nodes), Johnson-Cook, 227 m/s, 80 us, Courant
step every 10 increments.  One step = one complete solve of that model, from
t = 0 to 80 us, on a job whose state is already resident in HBM (jobs are
created before the timed region).

Metric: elem-cycles/s = n_elem x cycles / wall seconds.  ``value`` is the
mixed-precision path (FP32 strain / stress / radial return / forces, FP64
kinematics and accumulators); the plain FP64 path runs on the same model for
the speed-up and the error of the mixed result against it.

The roofline object is measured on the dominant kernel (the hex element
kernel) with HIP events recorded around its launches on the engine's stream
in one sampled cycle per graph during the timed region; its algorithmic bytes
per element-cycle are the compulsory state traffic of the layout (see
algo_bytes_per_elem); ``traffic`` is the rocprofv3 measurement committed under
profiles/ for this kernel and workload.

cpu_baseline: the FP64 CPU oracle (the restatement of the reference loop,
test infrastructure under oracle/) timed on this host, at 1 thread and on all
physical cores, over 20 cycles of the same model from 40 us (the plastic
phase), starting from the FP64 engine's state at that time.

Other workloads: --workload necking (round bar in tension, e.g. --n 40
--nz 208 for 998 400 hexes), sweep (batched seeded parameter sweep), serve
(serving loop with Poisson arrivals, p50/p99 latency).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dynela-v.-4.0_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def algo_bytes_per_elem(prec, nodes_per_elem):
    """Compulsory HBM bytes of one element-cycle in the element kernel.

    Per IP: the persistent state read and written once (FP64 path: stress 6,
    ep, dep, sy, gamma, gamma_cum, T, e_int, e_inel = 14 doubles each way;
    mixed path: stress 6 + dep, sy, gamma in FP32, ep, T, e_int, e_inel in
    FP64 = 68 B each way; its gamma_cum is ep / sqrt(2/3) at read-back) plus
    detJ0 read (8 B / 4 B).  Per
    element: connectivity 8 x 4 B, the per-element force vector written
    (24 x 8 B / 24 x 4 B).  Per node (x nodes/element): coordinates and the
    field0 increment read once (6 x 8 B).
    """
    if prec == "fp64":
        ip = 14 * 8 * 2 + 8
        force = 24 * 8
    else:
        ip = 68 * 2 + 4
        force = 24 * 4
    return 8 * ip + 8 * 4 + force + nodes_per_elem * 6 * 8


def make_model(prec, ctx, n=20, nz=100, t_end=80e-6, workload="taylor3d", pull=1.0):
    from dnlPython import samples
    if workload == "necking":
        return samples.necking_3d(n=n, nz=nz, t_end=t_end, pull=pull, precision=prec, track_tensors=False,
                                  context=ctx)
    return samples.taylor_3d(n=n, nz=nz, t_end=t_end, precision=prec, track_tensors=False, context=ctx)


def run_precision(prec, ctx, steps, warmup, n, nz, t_end, workload="taylor3d", pull=1.0):
    import dnlPython as dnl
    models = [make_model(prec, ctx, n, nz, t_end, workload, pull) for _ in range(steps + warmup)]
    for m in models:
        m.initSolve()
    for m in models[:warmup]:
        m.solve()
        print(f"[bench] {prec} warmup solve done", file=sys.stderr, flush=True)
    ctx.set_timing(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    for k, m in enumerate(models[warmup:]):
        m.solve()
        print(f"[bench] {prec} timed solve {k + 1}/{steps} done", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ctx.set_timing(False)
    el_ms, el_n = ctx.kernel_time("element")
    ne = models[0].getElementsNumber()
    nn = models[0].getNodesNumber()
    cycles = [m.getIncrement() for m in models[warmup:]]
    last = models[-1]
    disp = last.getNodalDisplacements()
    ep = last.getIntPointField(dnl.Field.plasticStrain)
    energies = last.getEnergies()       # kinetic, internal, plastic work, kinetic + internal
    mass = last.getNodalMasses()
    out = dict(wall=wall, cycles=cycles, ne=ne, nn=nn, elem_ms=el_ms, elem_launches=el_n,
               value=ne * sum(cycles) / wall, disp=disp, ep=ep, energies=energies, mass_total=float(mass.sum()),
               schedule=last.getSchedule())
    for m in models:
        m.close()
    return out


def run_sweep(prec, ctx, n_jobs, t_end, seed=20261203, max_cycles=None, perturb=0.0):
    """Batched parameter sweep: n_jobs seeded Taylor impacts (2D axisymmetric
    or 3D, 2k-60k elements log-uniform, Johnson-Cook A, B, n, C, m x U(0.8,
    1.2), 150-300 m/s) advanced together by one batched solve (one launch per
    phase and cycle for all jobs).  perturb: every impact speed times
    (1 + perturb), for the sensitivity of the solutions to their input."""
    import dnlPython as dnl
    from dnlPython import samples
    jobs = samples.sweep_params(n_jobs, seed=seed)
    for j in jobs:
        j["speed"] *= 1.0 + perturb
    models = [samples.build_sweep_job(j, t_end=t_end, precision=prec, context=ctx) for j in jobs]
    for m in models:
        m.initSolve()
    print(f"[bench] sweep {prec}: {n_jobs} jobs, {sum(m.getElementsNumber() for m in models)} elements",
          file=sys.stderr, flush=True)
    ctx.set_timing(True)
    ctx.reset_timing()
    t0 = time.perf_counter()
    statuses = dnl.batch_solve(models, raise_on_error=False, max_cycles=max_cycles)
    wall = time.perf_counter() - t0
    failed = [(i, s, models[i].getCurrentTime()) for i, s in enumerate(statuses) if s != 0]
    if failed:
        print(f"[bench] sweep {prec}: failed jobs (index, status, time): {failed}", file=sys.stderr, flush=True)
    capped = [(i, models[i].getCurrentTime(), models[i].getTimeStep()) for i, s in enumerate(statuses)
              if s == 0 and models[i].getCurrentTime() < t_end]
    if capped:
        ts = np.array([c[1] for c in capped])
        print(f"[bench] sweep {prec}: {len(capped)} jobs stopped by the cycle cap {max_cycles} at t = "
              f"{ts.min() * 1e6:.1f}-{ts.max() * 1e6:.1f} us (time step collapsed to "
              f"{min(c[2] for c in capped):.2e}-{max(c[2] for c in capped):.2e} s)", file=sys.stderr, flush=True)
    ctx.set_timing(False)
    el_ms, el_n = ctx.kernel_time("element")
    ec = sum(m.getElementsNumber() * m.getIncrement() for m in models)
    cyc = [m.getIncrement() for m in models]
    res = dict(wall=wall, value=ec / wall, elem_cycles=ec, jobs=n_jobs, failed=len(failed), capped=len(capped),
               elements=sum(m.getElementsNumber() for m in models), max_cycles=max(cyc),
               elem_ms=el_ms, elem_launches=el_n,
               done=[m.getCurrentTime() >= t_end for m in models], kinds=[j["kind"] for j in jobs],
               disp=[m.getNodalDisplacements() for m in models])
    for m in models:
        m.close()
    return res


def _wave_speed(mat):
    E, nu, rho = mat.youngModulus, mat.poissonRatio, mat.density
    return np.sqrt(E * (1.0 - nu) / (rho * (1.0 + nu) * (1.0 - 2.0 * nu)))


def estimate_work(model):
    """Rough elem-cycles of a job: n_elem x t_end / dt, dt from the Courant
    formula with the shortest element edge as length (for choosing the
    arrival rate only)."""
    a = model.descriptors()
    x = a.coords[:, a.conn]                       # [3, NE, nn]
    edges = np.linalg.norm(x - np.roll(x, 1, axis=2), axis=0).min()
    omega_s = 1.9818187003
    dt = 0.9 * omega_s * edges / (2.0 * _wave_speed(model._material))
    return a.conn.shape[0] * model._solver.endTime / dt


def run_serve(prec, ctx, n_jobs, util, rate, cycles_per_step, seed=20261203, cycle_budget=0):
    """Serving loop: Poisson arrivals of a seeded job mix (axisymmetric / 3D x
    impact / necking, 1k-500k elements log-uniform, t_end U(20, 100) us);
    jobs are submitted at their arrival times, admitted by device-memory
    footprint and advanced together (continuous batching).  Latency =
    completion - scheduled arrival."""
    import dnlPython as dnl
    from dnlPython import samples
    specs = samples.serve_params(n_jobs, seed=seed)
    models = [samples.build_serve_job(j, precision=prec, context=ctx) for j in specs]
    work = np.array([estimate_work(m) for m in models])
    if rate is None:
        # offered load = util x an assumed 1e9 elem-cycles/s service rate
        rate = util * 1.0e9 / work.mean()
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    print(f"[bench] serve {prec}: {n_jobs} jobs, {sum(m.getElementsNumber() for m in models)} elements, "
          f"rate {rate:.3g} jobs/s, estimated work {work.sum():.3g} elem-cycles", file=sys.stderr, flush=True)
    srv = dnl.Server(ctx, cycle_budget=cycle_budget)
    tickets = {}
    done = {}
    busy = 0.0
    nxt = 0
    step = cycles_per_step
    t0 = time.perf_counter()
    last_print = t0
    while len(done) < n_jobs:
        now = time.perf_counter() - t0
        while nxt < n_jobs and arrivals[nxt] <= now:
            tickets[srv.submit(models[nxt])] = nxt
            nxt += 1
            step = cycles_per_step
        if len(tickets) == len(done):
            if nxt < n_jobs:
                time.sleep(max(0.0, min(arrivals[nxt] - now, 0.05)))
            continue
        ts = time.perf_counter()
        running, queued = srv.step(step)
        el = time.perf_counter() - ts
        busy += el
        # steps of 10-40 ms: short enough to admit new arrivals promptly,
        # long enough that the host loop costs little
        if el < 0.010:
            step = min(step * 2, 8192)
        elif el > 0.040:
            step = max(cycles_per_step, step // 2)
        for t, i in list(tickets.items()):
            if i in done:
                continue
            state, _ = srv.poll(t)
            if state == dnl.Server.DONE:
                m, status = srv.take(t)
                done[i] = (time.perf_counter() - t0 - arrivals[i], status)
        if time.perf_counter() - last_print > 20.0:
            last_print = time.perf_counter()
            print(f"[bench] serve {prec}: {len(done)}/{n_jobs} done, {running} running, {queued} queued",
                  file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    lat = np.array([done[i][0] for i in range(n_jobs)])
    for i in np.argsort(-lat)[:8]:
        j = specs[i]
        print(f"[bench] serve job {i}: {j['kind']} {j['load']} {models[i].getElementsNumber()} elements, "
              f"t_end {j['t_end'] * 1e6:.1f} us, {models[i].getIncrement()} cycles, latency {lat[i]:.2f} s, "
              f"status {done[i][1]}", file=sys.stderr, flush=True)
    failed = [(i, done[i][1]) for i in range(n_jobs) if done[i][1] != 0]
    ec = sum(m.getElementsNumber() * m.getIncrement() for m in models)
    res = dict(wall=wall, value=ec / wall, elem_cycles=ec, p50=float(np.percentile(lat, 50)),
               p99=float(np.percentile(lat, 99)), busy=busy / wall, rate=rate, failed=failed,
               elements=sum(m.getElementsNumber() for m in models), span=float(arrivals[-1]))
    srv.close()
    for m in models:
        m.close()
    return res


def element_kernel(prec, schedule, dim=3, nn=8):
    """Name of the element-kernel instantiation a schedule launches (as the
    profiler prints it): <nodes, dim, block, axisymmetric, FP64 return,
    waves, table-law variant> for the mixed kernel, <nodes, dim, block,
    axisymmetric, network law, waves, tetrahedron> for the FP64 one."""
    b, w = schedule["elem_block"], schedule["waves"]
    if prec == "fp64":
        return f"k_elem_fp64<{nn}, {dim}, {b}, false, false, {w}, false>"
    return f"k_elem_mixed<{nn}, {dim}, {b}, false, {'true' if prec == 'mixed' else 'false'}, {w}, false>"


def measured_traffic(workload, kernel, n, nz):
    """HBM bytes per element-kernel launch measured offline with rocprofv3
    (FETCH_SIZE and WRITE_SIZE passes, scaled by a plain-copy calibration),
    when a summary exists for exactly this kernel and workload."""
    path = os.path.join(ROOT, "profiles", "element_kernels_hbm_traffic.json")
    if workload != "taylor3d" or (n, nz) != (20, 100) or not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    k = table["kernels"].get(kernel)
    return float(k["traffic_bytes_per_launch"]) if k else None


def host_cpu():
    """lscpu model, sockets, cores and nproc of this host."""
    info = {"nproc": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except Exception as e:  # noqa: BLE001 -- lscpu missing: report what os gives
        info["lscpu_error"] = str(e)
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return info


def physical_cores(info):
    """Physical cores this process may use (affinity / threads per core)."""
    usable = info.get("affinity") or info.get("nproc") or 1
    try:
        tpc = int(info.get("threads_per_core", 1))
    except ValueError:
        tpc = 1
    return max(1, usable // max(tpc, 1))


def cpu_baseline(n, nz, ctx, t_window=40e-6, window_cycles=20, workload="taylor3d", pull=1.0, threads=(1, None)):
    """The FP64 CPU oracle (restatement of the reference loop, test
    infrastructure) on the same model, timed on this host over a window of
    the plastic phase: the FP64 engine runs the job to t_window, its state
    goes to the oracle (checkpoint / restart), and the oracle advances
    window_cycles cycles from there, once per thread count (None: all
    physical cores)."""
    from oracle.bridge import OracleJob
    info = host_cpu()
    m = make_model("fp64", ctx, n, nz, 80e-6, workload, pull)
    m.solve(t_window)
    st = m.getState()
    ne = m.getElementsNumber()
    t_stop = st["time"] + window_cycles * st["timeStep"] * (1.0 - 1e-12)
    res = []
    for th in threads:
        th = physical_cores(info) if th is None else th
        job = OracleJob(m.descriptors(), threads=th)
        assert job.set_state(st) == 0
        t0 = time.perf_counter()
        status = job.solve(t_stop)
        wall = time.perf_counter() - t0
        cyc = job.scalars()[2] - st["increment"]
        res.append(dict(value=ne * cyc / wall, cores=th, wall_s=wall, cycles=cyc, status=status))
        print(f"[bench] cpu baseline {th} thread(s): {res[-1]['value']:.4g} elem-cycles/s, {cyc} cycles in "
              f"{wall:.1f} s", file=sys.stderr, flush=True)
    m.close()
    best = max(res, key=lambda r: r["value"])
    return dict(value=best["value"], unit="elem-cycles/s", cores=best["cores"], kind="port",
                sample=f"{best['cycles']} cycles x {ne} elements of the same model from t = {t_window * 1e6:g} us "
                       f"(plastic phase; state from the FP64 engine), FP64 oracle, element phases OpenMP-parallel",
                per_threads=[{k: r[k] for k in ("cores", "value", "wall_s", "cycles")} for r in res], host=info)


def copy_calibration(mib):
    """Plain device copy of `mib` MiB (reads and writes mib MiB), five times:
    the reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4}))


METRIC = "elem-cycles/s mixed vs FP64 at <=1e-5 rel disp error; p50/p99 solve latency"


def disp_rel_err(a, b):
    """max_i |a_i - b_i| / max_i |b_i| over node displacement vectors."""
    return float(np.linalg.norm(a - b, axis=0).max() / np.linalg.norm(b, axis=0).max())


def energy_balance(r, speed=227.0):
    """Closed system (the rigid wall does no work): |KE + IE - KE0| / KE0
    with KE0 = 1/2 M v0^2; the Chung-Hulbert numerical dissipation plus
    arithmetic error."""
    ke0 = 0.5 * r["mass_total"] * speed ** 2
    return float(abs(r["energies"][3] - ke0) / ke0)


def c1_legs(args, ctx, precs):
    """The 120 000-hex Taylor impact in each precision: the first is the
    headline (args.steps timed solves), the others args.side_steps."""
    res = {}
    for k, p in enumerate(precs):
        steps, warm = (args.steps, args.warmup) if k == 0 else (args.side_steps, min(args.warmup, 1))
        print(f"[bench] precision={p} steps={steps} warmup={warm}", file=sys.stderr, flush=True)
        res[p] = run_precision(p, ctx, steps, warm, args.n, args.nz, args.t_end, args.workload, args.pull)
        res[p]["steps"] = steps
        print(f"[bench] {p}: {res[p]['value']:.4g} elem-cycles/s, cycles={res[p]['cycles'][0]}, "
              f"wall={res[p]['wall']:.3f}s", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed solves of the headline precision")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="mixed,mixed-rr32,fp64",
                    help="precisions of the C1 legs; the first is the headline value")
    ap.add_argument("--side-steps", type=int, default=3, help="timed solves of the other precisions")
    ap.add_argument("--n", type=int, default=20)
    ap.add_argument("--nz", type=int, default=100)
    ap.add_argument("--t-end", type=float, default=80e-6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-window", type=float, default=40e-6, help="start of the CPU baseline's window (s)")
    ap.add_argument("--cpu-cycles", type=int, default=20, help="cycles of the CPU baseline's window")
    ap.add_argument("--no-batched", action="store_true", help="leave out the batched sweep leg")
    ap.add_argument("--no-serve", action="store_true", help="leave out the serving leg")
    ap.add_argument("--batched-jobs", type=int, default=256)
    ap.add_argument("--batched-cycle-cap", type=int, default=20000)
    ap.add_argument("--serve-jobs", type=int, default=24)
    ap.add_argument("--workload", default="taylor3d", choices=["taylor3d", "necking", "sweep", "serve"])
    ap.add_argument("--pull", type=float, default=1.0, help="necking: pulled-end speed (m/s)")
    ap.add_argument("--jobs", type=int, default=256, help="sweep / serve workloads: number of jobs")
    ap.add_argument("--max-cycles", type=int, default=None,
                    help="sweep: stop the batch after about this many cycles (jobs whose time step collapsed)")
    ap.add_argument("--no-batched-fp64", action="store_true",
                    help="leave out the FP64 run of the batched sweep (speedup and errors of the batched leg)")
    ap.add_argument("--no-sensitivity", dest="sensitivity", action="store_false",
                    help="sweep: leave out the FP64 rerun with perturbed impact speeds")
    ap.add_argument("--util", type=float, default=0.7, help="serve: offered load for the arrival rate")
    ap.add_argument("--rate", type=float, default=None, help="serve: arrival rate in jobs/s (overrides --util)")
    ap.add_argument("--step-cycles", type=int, default=32, help="serve: cycles per server step")
    ap.add_argument("--cycle-budget", type=int, default=200000,
                    help="serve: stop a job still short of t_end after this many increments (0: none)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain-copy reference kernel for the traffic counters")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(independent jobs would shard one job per GPU; one large mesh by element slabs)",
              file=sys.stderr)
        sys.exit(2)

    import dnlPython as dnl
    ctx = dnl.Context.default()
    precs = [p.strip() for p in args.precision.split(",") if p.strip()]
    dtype = lambda p: "f64" if p == "fp64" else "f32+f64"  # noqa: E731
    if args.workload == "serve":
        p = precs[0]
        r = run_serve(p, ctx, args.jobs, args.util, args.rate, args.step_cycles, cycle_budget=args.cycle_budget)
        line = {"metric": "served elem-cycles/s (continuous batching, Poisson arrivals)", "value": r["value"],
                "unit": "elem-cycles/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": r["wall"] * 1e3,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype(p),
                "data": "synthetic", "config": serve_config(r, args.jobs, p, args.step_cycles, args.cycle_budget),
                "latency_p50_s": r["p50"], "latency_p99_s": r["p99"], "busy_fraction": r["busy"],
                "failed_jobs": r["failed"]}
        print(json.dumps(line), flush=True)
        return
    if args.workload == "sweep":
        res = {}
        for p in precs:
            res[p] = run_sweep(p, ctx, args.jobs, args.t_end, max_cycles=args.max_cycles)
        r = res[precs[0]]
        line = {"metric": "batched elem-cycles/s (parameter sweep)", "value": r["value"], "unit": "elem-cycles/s",
                "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": r["wall"] * 1e3, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": dtype(precs[0]), "data": "synthetic",
                "config": sweep_config(r, precs[0], args.t_end, args.max_cycles)}
        for p in precs[1:]:
            line[f"value_{p}"] = res[p]["value"]
        if "mixed" in res and "fp64" in res:
            line["speedup_mixed_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
            line.update(sweep_errors(res["mixed"], res["fp64"], args, ctx))
        print(json.dumps(line), flush=True)
        return

    res = c1_legs(args, ctx, precs)
    head = precs[0]
    r = res[head]
    npe = r["nn"] / r["ne"]
    bytes_el = algo_bytes_per_elem(head, npe)
    avg_ms = r["elem_ms"] / max(r["elem_launches"], 1)
    kernel = element_kernel(head, r["schedule"])
    achieved = bytes_el * r["ne"] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    if args.workload == "taylor3d":
        wl = (f"Taylor impact 3D, {args.n}x{args.n}x{args.nz} O-grid hex8 ({r['ne']} elements, {r['nn']} nodes), "
              f"Johnson-Cook copper, 227 m/s, {args.t_end * 1e6:g} us")
    else:
        wl = (f"3D round bar necking, O-grid {args.n} x {args.nz} layers hex8 ({r['ne']} elements, {r['nn']} nodes), "
              f"Voce + linear law, 2 % taper, pulled at {args.pull:g} m/s, {args.t_end * 1e6:g} us")
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "elem-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype(head),
        "data": "synthetic",
        "config": {"workload": wl, "precision": head, "cycles_per_step": r["cycles"][0],
                   "schedule": r["schedule"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, kernel, args.n, args.nz),
                     "kernel": kernel,
                     "kernel_avg_ms": avg_ms, "algo_bytes_per_elem_cycle": bytes_el},
    }
    for p in precs[1:]:
        line[f"value_{p.replace('-', '_')}"] = res[p]["value"]
        line[f"steps_{p.replace('-', '_')}"] = res[p]["steps"]
    if "fp64" in res:
        d64, e64 = res["fp64"]["disp"], res["fp64"]["ep"]
        if "mixed" in res:
            line["speedup_mixed_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
        errs = {}
        for p in precs:
            if p == "fp64":
                continue
            errs[p] = {"disp": disp_rel_err(res[p]["disp"], d64),
                       "plastic_strain": float(np.abs(res[p]["ep"] - e64).max() / max(np.abs(e64).max(), 1e-300))}
        line["errors_vs_fp64"] = errs
        line["max_plastic_strain"] = float(np.abs(e64).max())
    if args.workload == "taylor3d":
        line["energy_balance_err"] = {p: energy_balance(res[p]) for p in precs}
    for p in precs:
        res[p].pop("disp", None)
        res[p].pop("ep", None)
    if args.workload == "taylor3d" and not args.no_batched:
        b = run_sweep(head, ctx, args.batched_jobs, 80e-6, max_cycles=args.batched_cycle_cap)
        line["batched"] = {"value": b["value"], "unit": "elem-cycles/s", "wall_s": b["wall"],
                           "config": sweep_config(b, head, 80e-6, args.batched_cycle_cap)}
        if head != "fp64" and not args.no_batched_fp64:
            # the same batch in FP64: the batched speedup and errors
            bf = run_sweep("fp64", ctx, args.batched_jobs, 80e-6, max_cycles=args.batched_cycle_cap)
            line["batched"]["value_fp64"] = bf["value"]
            line["batched"]["wall_s_fp64"] = bf["wall"]
            line["batched"][f"speedup_{head}_vs_fp64"] = b["value"] / bf["value"]
            line["batched"].update(sweep_errors(b, bf, args, ctx, jobs=args.batched_jobs, cap=args.batched_cycle_cap))
    if args.workload == "taylor3d" and not args.no_serve:
        sv = run_serve(head, ctx, args.serve_jobs, 0.7, None, 32, cycle_budget=args.cycle_budget)
        line["serving"] = {"value": sv["value"], "unit": "elem-cycles/s", "latency_p50_s": sv["p50"],
                           "latency_p99_s": sv["p99"], "busy_fraction": sv["busy"], "wall_s": sv["wall"],
                           "failed_jobs": sv["failed"],
                           "config": serve_config(sv, args.serve_jobs, head, 32, args.cycle_budget)}
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args.n, args.nz, ctx, t_window=min(args.cpu_window, 0.5 * args.t_end),
                                            window_cycles=args.cpu_cycles, workload=args.workload, pull=args.pull)
    print(json.dumps(line), flush=True)


def sweep_errors(mx, fp, args, ctx, jobs=None, cap=None):
    """Mixed vs FP64 displacement errors of the sweep's jobs that both runs
    finished (overall and per kind) and, with args.sensitivity, the FP64
    solutions' own sensitivity: the same FP64 sweep with every impact speed
    times (1 + 2^-24), the FP32 unit roundoff -- how far the jobs' answers
    move under an input change of the size that mixed precision makes at
    every step."""
    out = {}
    both = [i for i in range(len(fp["done"])) if mx["done"][i] and fp["done"][i]]
    errs = [disp_rel_err(mx["disp"][i], fp["disp"][i]) for i in both]
    if not errs:
        return out
    out["mixed_disp_err_vs_fp64"] = {"max": max(errs), "median": float(np.median(errs)), "jobs_compared": len(errs)}
    for kind in ("axi", "3d"):
        ek = [e for e, i in zip(errs, both) if fp["kinds"][i] == kind]
        if ek:
            out["mixed_disp_err_vs_fp64"][f"max_{kind}"] = max(ek)
            out["mixed_disp_err_vs_fp64"][f"jobs_{kind}"] = len(ek)
    if not args.sensitivity:
        return out
    pr = run_sweep("fp64", ctx, jobs or args.jobs, 80e-6 if jobs else args.t_end,
                   max_cycles=cap if jobs else args.max_cycles, perturb=2.0 ** -24)
    sens = [disp_rel_err(pr["disp"][i], fp["disp"][i]) if pr["done"][i] else float("nan") for i in both]
    if not any(s == s for s in sens):
        return out
    ratio = [e / s for e, s in zip(errs, sens) if s == s and s > 0]
    out["fp64_sensitivity"] = {
        "perturbation": "impact speed x (1 + 2^-24)", "max": float(np.nanmax(sens)), "median": float(np.nanmedian(sens)),
        "mixed_err_over_sensitivity_max": max(ratio) if ratio else None,
        "mixed_err_over_sensitivity_median": float(np.median(ratio)) if ratio else None}
    for kind in ("axi", "3d"):
        sk = [s for s, i in zip(sens, both) if fp["kinds"][i] == kind and s == s]
        if sk:
            out["fp64_sensitivity"][f"max_{kind}"] = max(sk)
            out["fp64_sensitivity"][f"median_{kind}"] = float(np.median(sk))
    return out


def sweep_config(r, prec, t_end, cap):
    return {"workload": f"batched sweep of {r['jobs']} seeded Taylor impacts (2D axisymmetric / 3D hex, "
                        f"{r['elements']} elements in total, 2k-60k each, Johnson-Cook A, B, n, C, m x U(0.8, 1.2), "
                        f"150-300 m/s), {t_end * 1e6:g} us",
            "precision": prec, "max_cycles": r["max_cycles"], "failed_jobs": r["failed"], "cycle_cap": cap,
            "jobs_stopped_by_cap": r["capped"], "elem_cycles": r["elem_cycles"]}


def serve_config(r, n_jobs, prec, step_cycles, budget):
    return {"workload": f"serving mix of {n_jobs} jobs (axisymmetric / 3D x impact / necking, {r['elements']} "
                        f"elements in total, 1k-500k each, t_end 20-100 us), Poisson arrivals",
            "precision": prec, "arrival_rate_jobs_per_s": r["rate"], "cycles_per_step": step_cycles,
            "cycle_budget": budget, "elem_cycles": r["elem_cycles"]}


if __name__ == "__main__":
    main()
