"""CPU checks of the drop-in boundary: libdnlg.so loads and exports every
entry point declared in include/dnlg.h (no compute call without a GPU), the
ctypes structures match the header, and the dnlPython model-building layer
produces the descriptors the reference semantics require."""
import ctypes
import os
import re

import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import _lib, samples

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dnlg.h")


def declared_entry_points():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dnlg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_entry_points()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert set(declared_entry_points()) == set(_lib.SIGNATURES)


def test_abi_version_and_footprint_without_gpu():
    lib = _lib.load()
    assert lib.dnlg_abi_version() == 2
    b64 = lib.dnlg_footprint_bytes(_lib.EL_HEX8N_3D, 125341, 120000, _lib.PREC_FP64, 0)
    bmx = lib.dnlg_footprint_bytes(_lib.EL_HEX8N_3D, 125341, 120000, _lib.PREC_MIXED, 0)
    assert b64 > 0 and bmx > 0 and bmx < b64 + 120000 * 8 * 4 * 24
    assert lib.dnlg_footprint_bytes(99, 10, 10, 0, 0) == -1


def test_struct_layouts_match_header():
    # offsets follow from the C declarations (natural alignment, LP64)
    assert ctypes.sizeof(_lib.Mesh) == 40
    assert _lib.Mesh.coords.offset == 24
    assert ctypes.sizeof(_lib.MaterialDesc) == 6 * 8 + 8 + 64 + 8 + 8
    assert _lib.MaterialDesc.table.offset == 6 * 8 + 8 + 64
    assert ctypes.sizeof(_lib.BoundaryDesc) == 8 + 24 + 16
    assert ctypes.sizeof(_lib.Bcs) == 8 + 5 * 8
    assert ctypes.sizeof(_lib.SolverParams) == 4 * 8 + 6 * 4


def test_context_creation_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(dnl.DnlgError):
        dnl.Context(0)


def test_taylor_axi_descriptors_follow_the_reference_sample():
    """Samples/Taylor/Taylor-Axi/Taylor.py: node numbering, CCW connectivity,
    bottom row restrained in y, axis restrained in x, initial speed on all."""
    m = samples.taylor_axi(nr=5, nz=4)
    a = m.descriptors()
    assert a.coords.shape == (3, 30) and a.conn.shape == (20, 4)
    np.testing.assert_array_equal(a.conn[0], [0, 1, 7, 6])
    # every quad counter-clockwise in (r, z)
    x, y = a.coords[0][a.conn], a.coords[1][a.conn]
    area = 0.5 * np.sum(x * np.roll(y, -1, 1) - np.roll(x, -1, 1) * y, axis=1)
    assert np.all(area > 0)
    b = list(a.boundaries)[: a.bcs.n_boundaries]
    kinds = [(bb.kind, tuple(bb.value)) for bb in b]
    assert (_lib.BC_SPEED, (0.0, -227.0, 0.0)) in kinds
    # constant lists: node 1 (origin) has bottom then axis, in attach order
    ptr, idx = a.const_ptr, a.const_idx
    first = [kinds[i] for i in idx[ptr[0]:ptr[1]]]
    assert first == [(_lib.BC_RESTRAIN, (0.0, 1.0, 1.0)), (_lib.BC_RESTRAIN, (1.0, 0.0, 1.0))]
    # a node on the axis above the bottom: axis only
    n7 = [kinds[i] for i in idx[ptr[6]:ptr[7]]]
    assert n7 == [(_lib.BC_RESTRAIN, (1.0, 0.0, 1.0))]
    # a free interior node: none
    assert ptr[8] == ptr[9]
    assert a.init_ptr[-1] == 30


def test_nodes_sorted_by_number_and_set_ranges():
    m = dnl.DynELA("t")
    m.createNode(30, 1.0, 0.0, 0.0)
    m.createNode(10, 0.0, 0.0, 0.0)
    m.createNode(20, 1.0, 1.0, 0.0)
    m.createNode(40, 0.0, 1.0, 0.0)
    m.setDefaultElement(dnl.Element.ElQua4NAx)
    m.createElement(1, 10, 30, 20, 40)
    ns = dnl.NodeSet("ns")
    m.add(ns, 10, 40, 10)   # Model::add(set, start, end, increment): 10, 20, 30, 40
    assert ns._numbers == [10, 20, 30, 40]
    ns2 = dnl.NodeSet("one")
    m.add(ns2, 30)
    assert ns2._numbers == [30]
    es = dnl.ElementSet("es")
    m.add(es, 1)
    law = dnl.JohnsonCookLaw()
    law.setParameters(1, 1, 0, 1, 1, 1, 1000, 20)
    mat = dnl.Material("m")
    mat.setHardeningLaw(law)
    mat.youngModulus, mat.poissonRatio, mat.density, mat.heatCapacity = 1e9, 0.3, 1000.0, 500.0
    m.add(mat, es)
    s = dnl.Explicit("s")
    s.setTimes(0, 1e-6)
    m.add(s)
    a = m.descriptors()
    np.testing.assert_array_equal(a.coords[0], [0.0, 1.0, 1.0, 0.0])   # nodes 10, 20, 30, 40
    np.testing.assert_array_equal(a.conn[0], [0, 2, 1, 3])


def test_model_errors_are_exceptions():
    m = dnl.DynELA("bad")
    with pytest.raises(dnl.DnlgError):
        m.createElement(1, 1, 2, 3, 4)            # no default element
    m.setDefaultElement(dnl.Element.ElQua4NAx)
    with pytest.raises(dnl.DnlgError):
        m.createElement(1, 1, 2, 3)                # wrong node count
    with pytest.raises(dnl.DnlgError):
        m.setDefaultElement(dnl.Element.ElTet4N3D)
        m.createElement(2, 1, 2, 3, 4)             # element family outside the engine
    with pytest.raises(dnl.DnlgError):
        dnl.Explicit("s").setDissipation(1.5)      # Explicit.C:112-123 range check
    m2 = samples.taylor_axi(nr=2, nz=2)
    m2._node_num.append(np.array([1]))
    m2._node_xyz.append(np.zeros((1, 3)))
    with pytest.raises(dnl.DnlgError):
        m2.descriptors()                            # duplicate node number


def test_ogrid_cross_section_is_valid():
    pts, quads = samples.ogrid_disc(6, 3)
    assert quads.shape[0] == 6 * 6 + 4 * 6 * 3
    x, z = pts[quads, 0], pts[quads, 1]
    area = 0.5 * np.sum(x * np.roll(z, -1, 1) - np.roll(x, -1, 1) * z, axis=1)
    assert np.all(area > 0)
    # the boundary is the regular 24-gon inscribed in the unit circle
    n = 4 * 6
    assert area.sum() == pytest.approx(n / 2 * np.sin(2 * np.pi / n), rel=1e-13)
    # every interior edge shared by exactly two cells, boundary edges on the circle
    edges = {}
    for q in quads:
        for k in range(4):
            e = tuple(sorted((q[k], q[(k + 1) % 4])))
            edges[e] = edges.get(e, 0) + 1
    boundary = [e for e, c in edges.items() if c == 1]
    assert all(c in (1, 2) for c in edges.values())
    r = np.hypot(pts[:, 0], pts[:, 1])
    assert np.allclose(r[np.array(boundary).ravel()], 1.0)


def test_sweep_parameters_are_seeded_and_in_range():
    a = samples.sweep_params(64)
    b = samples.sweep_params(64)
    assert a == b
    for j in a:
        assert 150.0 <= j["speed"] <= 300.0
        for k in ("A", "B", "n", "C", "m"):
            assert 0.8 * samples.COPPER[k] <= j["material"][k] <= 1.2 * samples.COPPER[k]


def test_serve_mix_is_seeded_and_in_range():
    a = samples.serve_params(48)
    assert a == samples.serve_params(48)
    kinds = {(j["kind"], j["load"]) for j in a}
    assert kinds == {("axi", "impact"), ("axi", "necking"), ("3d", "impact"), ("3d", "necking")}
    for j in a:
        assert 20e-6 <= j["t_end"] <= 100e-6
        m = samples.build_serve_job(j, precision="fp64")
        ne = m.getElementsNumber()
        assert 700 <= ne <= 700000


def test_necking_axi_descriptors():
    """Axisymmetric bar: CCW quads, bottom row axial restraint, axis radial
    restraint, top row constant pull speed; mid-length radius reduced by the
    taper."""
    m = samples.necking_axi(nr=4, nz=8, pull=2.0)
    a = m.descriptors()
    x, y = a.coords[0][a.conn], a.coords[1][a.conn]
    area = 0.5 * np.sum(x * np.roll(y, -1, 1) - np.roll(x, -1, 1) * y, axis=1)
    assert np.all(area > 0)
    r_outer = a.coords[0].reshape(9, 5)[:, -1]
    assert r_outer[4] == pytest.approx(6.413e-3 * 0.98, rel=1e-12)
    assert r_outer[0] == pytest.approx(6.413e-3, rel=1e-12)
    kinds = [(b.kind, tuple(b.value)) for b in list(a.boundaries)[: a.bcs.n_boundaries]]
    ptr, idx = a.const_ptr, a.const_idx
    top = [kinds[i] for i in idx[ptr[8 * 5 + 2]:ptr[8 * 5 + 3]]]
    assert top == [(_lib.BC_SPEED, (0.0, 2.0, 0.0))]


def test_context_close_releases_its_models_jobs():
    """Closing a context first releases the engine jobs of its models (a job
    may not outlive its context); the models' later close is a no-op."""

    class FakeLib:
        def __init__(self):
            self.calls = []

        def dnlg_job_destroy(self, job):
            self.calls.append(("job", job))

        def dnlg_ctx_destroy(self, h):
            self.calls.append(("ctx", h))

    ctx = object.__new__(dnl.Context)
    ctx.lib, ctx.handle = FakeLib(), ctypes.c_void_p(1)
    import weakref
    ctx._models = weakref.WeakSet()
    m = dnl.DynELA("m", context=ctx)
    m._job = ctypes.c_void_p(7)
    ctx._models.add(m)
    ctx.close()
    assert [c[0] for c in ctx.lib.calls] == ["job", "ctx"]
    assert m._job is None
    m.close()
    assert len(ctx.lib.calls) == 2


def test_header_constants_match_the_binding():
    """Every integer #define DNLG_X of include/dnlg.h that the binding names
    X has the same value there."""
    src = open(HEADER).read()
    defines = dict((k, int(v)) for k, v in re.findall(r"#define DNLG_([A-Z0-9_]+)\s+\(?(-?\d+)\)?", src))
    checked = 0
    for name, value in defines.items():
        if hasattr(_lib, name):
            assert getattr(_lib, name) == value, (name, getattr(_lib, name), value)
            checked += 1
    assert checked >= 25
    assert defines["LAW_ANN2SIG"] == _lib.LAW_ANN2SIG and defines["ANN_MAX_HIDDEN"] == _lib.ANN_MAX_HIDDEN
