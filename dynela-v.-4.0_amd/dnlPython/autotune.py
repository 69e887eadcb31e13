"""Schedule table of the engine: measure (``tune``) and load (``load``).

``tune`` builds a representative Taylor impact for every element family
(hexahedra, axisymmetric quads), precision and job-size bucket, runs it into
its plastic phase and lets the engine's autotuner (dnlg_job_autotune)
measure the candidate launch schedules on it: threads per element-kernel
block and waves per SIMD, the corrector's slot batching, cycles per graph.
The winners become the rules of ``schedules.json`` next to the library, which
``load`` hands to a context (Context does it when it is created).  A
schedule never changes results (the engine's kernels compute every element
and node the same way under all of them), only the time.

The reference has no counterpart: its element loops take fixed OpenMP
chunks (Parallel.C:78-113).

    python -m dnlPython.autotune            # on the GPU; writes schedules.json
"""
import datetime
import json
import os
import sys

from . import _lib

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "schedules.json")
PRECISIONS = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED, "mixed-rr32": _lib.PREC_MIXED_RR32}
# job-size buckets [lo, hi) in elements (hi 0: unbounded) and the
# representative model of each: (family, builder arguments)
BUCKETS = [(0, 10000), (10000, 100000), (100000, 0)]
REPRESENTATIVE = {
    ("hex", 0): dict(n=6, nz=30),      # 3 240 hexes
    ("hex", 1): dict(n=10, nz=100),    # 30 000
    ("hex", 2): dict(n=20, nz=100),    # 120 000
    ("axi", 0): dict(nr=10, nz=300),   # 3 000 quads
    ("axi", 1): dict(nr=30, nz=1000),  # 30 000
    ("axi", 2): dict(nr=60, nz=2000),  # 120 000
}
# element-cycles per second used to size a trial to about 20 ms
_RATE = {"fp64": 2.5e8, "mixed": 1.8e9, "mixed-rr32": 2.5e9}


def _model(family, bucket, precision, ctx):
    from . import samples
    kw = dict(REPRESENTATIVE[(family, bucket)], precision=precision, track_tensors=False, context=ctx,
              t_end=2e-3)
    return samples.taylor_3d(**kw) if family == "hex" else samples.taylor_axi(**kw)


def tune(ctx=None, out=TABLE, families=("hex", "axi"), precisions=("mixed", "mixed-rr32", "fp64"),
         buckets=(0, 1, 2), rounds=3, log=sys.stderr):
    """Measure and write the table; returns it."""
    from . import Context
    ctx = ctx or Context.default()
    ctx.clear_schedules()
    rules = []
    for family in families:
        for prec in precisions:
            for b in buckets:
                m = _model(family, b, prec, ctx)
                ne = m.getElementsNumber()
                m.solve(30e-6)  # into the plastic phase of the impact
                per_cycle = max(20e-6, ne / _RATE[prec])
                cycles = int(min(1000, max(40, 0.02 / per_cycle)))
                sched, ms, ms0 = m.autotune(cycles_per_trial=cycles, rounds=rounds)
                lo, hi = BUCKETS[b]
                rule = dict(elem_type=m._elem_type, precision=PRECISIONS[prec], min_elems=lo, max_elems=hi,
                            schedule=sched, ms_per_cycle=ms, default_ms_per_cycle=ms0,
                            model=f"Taylor impact {family} {ne} elements, from 30 us", cycles_per_trial=cycles,
                            rounds=rounds)
                rules.append(rule)
                print(f"[autotune] {family} {prec} {ne} elements: {sched} {ms * 1e3:.1f} us/cycle "
                      f"(default {ms0 * 1e3:.1f})", file=log, flush=True)
                m.close()
    table = {"device": "MI355X (gfx950)", "created": datetime.datetime.utcnow().isoformat(timespec="seconds"),
             "rules": rules}
    if out:
        with open(out, "w") as f:
            json.dump(table, f, indent=1)
    load(ctx, table)
    return table


def load(ctx, table=TABLE):
    """Hand the rules of a table (a dict or a path) to a context; returns
    the number of rules (0 when there is no table)."""
    if isinstance(table, str):
        if not os.path.exists(table):
            return 0
        with open(table) as f:
            table = json.load(f)
    ctx.clear_schedules()
    for r in table["rules"]:
        ctx.add_schedule(r["elem_type"], r["precision"], r["min_elems"], r["max_elems"], **r["schedule"])
    return len(table["rules"])


if __name__ == "__main__":
    tune()
