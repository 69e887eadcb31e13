"""Launch schedules (no GPU): the candidates the kernels are built for, and
the committed schedule table (dynela-v.-4.0_amd/schedules.json, written by
dnlPython.autotune on the GPU) names only those."""
import ctypes
import json
import os

import pytest

from dnlPython import _lib
from dnlPython.autotune import BUCKETS, TABLE

HEX, AXI, PLANE = _lib.EL_HEX8N_3D, _lib.EL_QUA4N_AX, _lib.EL_QUA4N_2D
JC, ANN = _lib.LAW_JOHNSON_COOK, _lib.LAW_ANN2SIG


def valid(t, prec, law=JC, **kw):
    s = _lib.Schedule(**kw)
    return _lib.load().dnlg_schedule_valid(t, prec, law, ctypes.byref(s)) == 1


def test_candidates_are_built():
    for t, blocks in ((HEX, (64, 128, 256)), (AXI, (128, 256, 512))):
        for b in blocks:
            for prec, alt in ((_lib.PREC_MIXED, 5), (_lib.PREC_MIXED_RR32, 5), (_lib.PREC_FP64, 2)):
                assert valid(t, prec, elem_block=b)
                assert valid(t, prec, elem_block=b, waves=alt)
    assert valid(HEX, _lib.PREC_FP64, waves=3) and valid(HEX, _lib.PREC_MIXED, waves=0)
    for c in (2, 4, 8):
        assert valid(HEX, _lib.PREC_MIXED, corr_batch=c)
    assert valid(HEX, _lib.PREC_FP64, corr_batch=4) and not valid(HEX, _lib.PREC_FP64, corr_batch=8)


def test_other_schedules_are_refused():
    assert not valid(HEX, _lib.PREC_MIXED, elem_block=512)      # a quad block size
    assert not valid(AXI, _lib.PREC_MIXED, elem_block=64)
    assert not valid(HEX, _lib.PREC_MIXED, waves=4)             # not built: the natural occupancy
    assert not valid(HEX, _lib.PREC_MIXED, corr_batch=3)
    assert not valid(HEX, _lib.PREC_MIXED, graph_cycles=1)
    # the plane-strain quad and the network law run their defaults only
    assert valid(PLANE, _lib.PREC_MIXED) and not valid(PLANE, _lib.PREC_MIXED, elem_block=128)
    assert valid(HEX, _lib.PREC_FP64, ANN) and not valid(HEX, _lib.PREC_FP64, ANN, waves=2)


@pytest.mark.skipif(not os.path.exists(TABLE), reason="no schedule table measured yet")
def test_committed_table_names_built_schedules():
    with open(TABLE) as f:
        table = json.load(f)
    seen = set()
    for r in table["rules"]:
        assert valid(r["elem_type"], r["precision"], **r["schedule"]), r
        assert (r["min_elems"], r["max_elems"]) in BUCKETS
        assert r["ms_per_cycle"] > 0 and r["ms_per_cycle"] <= r["default_ms_per_cycle"] * 1.05
        seen.add((r["elem_type"], r["precision"], r["min_elems"]))
    for t in (HEX, AXI):
        for p in (_lib.PREC_FP64, _lib.PREC_MIXED, _lib.PREC_MIXED_RR32):
            for lo, _ in BUCKETS:
                assert (t, p, lo) in seen
