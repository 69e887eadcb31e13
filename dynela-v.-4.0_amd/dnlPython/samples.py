"""Model builders for the benchmark workloads (bulk numpy ingestion).

Each builder returns a ready ``DynELA`` model (not yet solved).  Units are
SI (m, kg, s, Pa) with temperatures in degrees Celsius for the Johnson-Cook
law.  Node and element numbering follows the reference samples: nodes
row-major from 1, elements from 1 (Samples/Taylor/Taylor-Axi/Taylor.py:43-67).

Workloads
  taylor_axi   2D axisymmetric Taylor impact, 4-node quads (10 x 50)
  taylor_3d    3D Taylor impact, full cylinder of 8-node hexes (O-grid
               cross-section of 3 n^2 cells, nz layers; n=20, nz=100 gives
               120 000 hexes)
  necking_3d   3D round bar in tension with a 2 % mid-length taper,
               Voce + linear hardening, one end fixed, the other pulled
  necking_axi  the same bar with 4-node axisymmetric quads
  necking_plane the same half-plane as a plane-strain plate (ElQua4N2D)
  taylor_tet   3D impact of a square copper bar of 4-node tetrahedra
               (ElTet4N3D, six per cell), power-iteration time step
  sweep_params seeded parameter sweep around the copper impact
  serve_params seeded serving mix (axi / 3D x impact / necking)
"""
import numpy as np

from . import (BoundaryRestrain, BoundarySpeed, DynELA, Element, ElementSet, Explicit, JohnsonCookLaw,
               Material, NodeSet, VoceLinearLaw)

# Copper-like Johnson-Cook set of the impact workloads; Tm, T0, Cp and eta
# are not given with the impact configuration and are copper handbook values.
COPPER = dict(young=124e9, poisson=0.34, density=8930.0, A=90e6, B=292e6, C=0.025, n=0.31, m=1.09,
              depsp0=1.0, Tm=1083.0, T0=20.0, heatCapacity=383.0, taylorQuinney=0.9)

# Necking bar: Voce + linear law sigma_y = 80 + 300 (1 - exp(-16.93 p)) + 129.24 p MPa;
# elastic and thermal constants are steel-like values chosen for this workload.
NECKING = dict(young=206e9, poisson=0.3, density=7830.0, A=80e6, Q=300e6, b=16.93, H=129.24e6,
               T0=20.0, heatCapacity=460.0, taylorQuinney=0.9)


def _jc_material(p):
    law = JohnsonCookLaw("JC")
    law.setParameters(p["A"], p["B"], p["C"], p["n"], p["m"], p["depsp0"], p["Tm"], p["T0"])
    mat = Material("copper")
    mat.setHardeningLaw(law)
    mat.youngModulus = p["young"]
    mat.poissonRatio = p["poisson"]
    mat.density = p["density"]
    mat.heatCapacity = p["heatCapacity"]
    mat.taylorQuinney = p["taylorQuinney"]
    mat.T0 = p["T0"]
    return mat


def _solver(t_end, precision, track_tensors, freq=10, dt_method=None):
    s = Explicit("Solver")
    s.setTimes(0.0, t_end)
    s.setComputeTimeStepFrequency(freq)
    if dt_method is not None:
        s.setTimeStepMethod(dt_method)
    s.setPrecision(precision)
    s.setTrackTensors(track_tensors)
    return s


def _set(numbers, name):
    ns = NodeSet(name)
    ns._numbers.extend(int(v) for v in numbers)
    return ns


def taylor_axi(nr=10, nz=50, radius=3.2e-3, length=32.4e-3, speed=227.0, t_end=80e-6, material=None,
               precision="fp64", track_tensors=True, context=None, freq=10, dt_method=None):
    """Taylor-Axi impact (Samples/Taylor/Taylor-Axi/Taylor.py): r = x, axis = y,
    bottom row restrained in y, axis nodes restrained in x (and z),
    initial axial speed -speed on every node."""
    p = dict(COPPER)
    if material:
        p.update(material)
    model = DynELA("Taylor-Axi", context=context)
    i, j = np.meshgrid(np.arange(nr + 1), np.arange(nz + 1))
    nums = (j * (nr + 1) + i + 1).ravel()
    xyz = np.stack([(i * (radius / nr)).ravel(), (j * (length / nz)).ravel(), np.zeros(nums.size)], axis=1)
    model.createNodes(nums, xyz)
    model.setDefaultElement(Element.ElQua4NAx)
    ei, ej = np.meshgrid(np.arange(nr), np.arange(nz))
    n1 = ei + ej * (nr + 1) + 1
    conn = np.stack([n1, n1 + 1, ei + (ej + 1) * (nr + 1) + 2, ei + (ej + 1) * (nr + 1) + 1], axis=-1)
    model.createElements(np.arange(1, nr * nz + 1), conn.reshape(-1, 4))
    allES = ElementSet("ES_All")
    model.add(allES, 1, nr * nz)
    model.add(_jc_material(p), allES)
    bottom = BoundaryRestrain("BC_bottom")
    bottom.setValue(0, 1, 1)
    model.attachConstantBC(bottom, _set(range(1, nr + 2), "NS_Bottom"))
    axis = BoundaryRestrain("BC_axis")
    axis.setValue(1, 0, 1)
    model.attachConstantBC(axis, _set(range(1, nums.size + 1, nr + 1), "NS_Axis"))
    v = BoundarySpeed("BC_speed")
    v.setValue(0, -speed, 0)
    model.attachInitialBC(v, _set(nums, "NS_All"))
    model.add(_solver(t_end, precision, track_tensors, freq, dt_method))
    return model


def ogrid_disc(m, k, s=0.42):
    """O-grid ("butterfly") quad mesh of the unit disc: an m x m square core of
    half-width s plus four blocks of m x k cells joining the core edges to the
    circle.  Returns points [P, 2] (x, z) and counter-clockwise quads [Q, 4]
    (indices into points); m*m + 4*m*k cells."""
    pts, quads = [], []
    u = np.linspace(-1.0, 1.0, m + 1)
    # core
    a, b = np.meshgrid(u, u, indexing="ij")
    core = np.stack([s * a, s * b], axis=-1)
    base = 0
    pts.append(core.reshape(-1, 2))
    ia, ib = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    q = lambda i, j: base + i * (m + 1) + j  # noqa: E731
    quads.append(np.stack([q(ia, ib), q(ia + 1, ib), q(ia + 1, ib + 1), q(ia, ib + 1)], -1).reshape(-1, 4))
    base += (m + 1) * (m + 1)
    # four radial blocks: edge parameter t in [-1, 1], radial j in [0, k]
    w = np.linspace(0.0, 1.0, k + 1)
    for rot in range(4):
        ang = rot * np.pi / 2
        c, sn = np.cos(ang), np.sin(ang)
        t, jj = np.meshgrid(u, w, indexing="ij")
        pin = np.stack([np.full_like(t, s), s * t], -1)
        th = (np.pi / 4) * t
        pout = np.stack([np.cos(th), np.sin(th)], -1)
        p = (1 - jj)[..., None] * pin + jj[..., None] * pout
        p = np.stack([c * p[..., 0] - sn * p[..., 1], sn * p[..., 0] + c * p[..., 1]], -1)
        pts.append(p.reshape(-1, 2))
        it, ij = np.meshgrid(np.arange(m), np.arange(k), indexing="ij")
        q = lambda i, j, b0=base: b0 + i * (k + 1) + j  # noqa: E731
        quads.append(np.stack([q(it, ij), q(it, ij + 1), q(it + 1, ij + 1), q(it + 1, ij)], -1).reshape(-1, 4))
        base += (m + 1) * (k + 1)
    pts = np.concatenate(pts)
    quads = np.concatenate(quads)
    # merge coincident points (core edges and block diagonals)
    key = np.round(pts / (1e-9 + 0.25 / max(m, k)) * 1e6).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(order.size)
    pts = pts[first[order]]
    quads = remap[inv.ravel()[quads]]
    # counter-clockwise orientation in the (x, z) plane
    x, z = pts[quads, 0], pts[quads, 1]
    area = 0.5 * np.sum(x * np.roll(z, -1, 1) - np.roll(x, -1, 1) * z, axis=1)
    quads[area < 0] = quads[area < 0][:, ::-1]
    return pts, quads


def _hex_cylinder(n, nz, radius, ys, k=None):
    """Full-cylinder hex mesh, axis y, O-grid cross-section (n x n core cells
    plus 4 x n x k ring cells).  ys: nz+1 axial levels; radius may be an
    array over levels.  Returns numbers, xyz [N,3], conn [NE,8] (1-based node
    numbers) and the node count of one level."""
    k = max(1, n // 2) if k is None else k
    pts, quads = ogrid_disc(n, k)
    nl = pts.shape[0]
    r = np.broadcast_to(np.asarray(radius, dtype=np.float64), (nz + 1,))
    xyz = np.zeros((nz + 1, nl, 3))
    xyz[..., 0] = r[:, None] * pts[None, :, 0]
    xyz[..., 1] = np.asarray(ys)[:, None]
    xyz[..., 2] = r[:, None] * pts[None, :, 1]
    nums = np.arange(1, (nz + 1) * nl + 1)
    lev = np.arange(nz)[:, None, None] * nl
    # quads are counter-clockwise in (x, z); with the axis along y the
    # right-handed hex order (detJ > 0) walks each face clockwise in (x, z)
    bot = quads[:, ::-1][None] + lev + 1
    conn = np.concatenate([bot, bot + nl], axis=-1).reshape(-1, 8)
    return nums, xyz.reshape(-1, 3), conn, nl


def taylor_3d(n=20, nz=100, radius=3.2e-3, length=32.4e-3, speed=227.0, t_end=80e-6, material=None,
              precision="fp64", track_tensors=True, context=None, freq=10, dt_method=None):
    """3D Taylor impact of a full cylinder (no symmetry planes): O-grid
    cross-section of n*n + 4*n*(n/2) hexes, nz layers (n=20, nz=100 ->
    120 000 hexes); bottom face restrained along the axis (y), initial speed
    -speed on every node."""
    p = dict(COPPER)
    if material:
        p.update(material)
    model = DynELA("Taylor-3D", context=context)
    nums, xyz, conn, nl = _hex_cylinder(n, nz, radius, np.linspace(0.0, length, nz + 1))
    model.createNodes(nums, xyz)
    model.setDefaultElement(Element.ElHex8N3D)
    ne = conn.shape[0]
    model.createElements(np.arange(1, ne + 1), conn)
    allES = ElementSet("ES_All")
    model.add(allES, 1, ne)
    model.add(_jc_material(p), allES)
    bottom = BoundaryRestrain("BC_bottom")
    bottom.setValue(0, 1, 0)
    model.attachConstantBC(bottom, _set(range(1, nl + 1), "NS_Bottom"))
    v = BoundarySpeed("BC_speed")
    v.setValue(0, -speed, 0)
    model.attachInitialBC(v, _set(nums, "NS_All"))
    model.add(_solver(t_end, precision, track_tensors, freq, dt_method))
    return model


def _tet_box(nx, ny, nz, lx, ly, lz):
    """Box [0, lx] x [0, ly] x [0, lz] of nx ny nz cells, each split into six
    tetrahedra around its diagonal from corner (0, 0, 0) to (1, 1, 1); the
    same split in every cell makes shared faces match.  Node order of each
    tetrahedron gives a positive volume (ElTet4N3D's node 0 at the origin of
    its local axes, nodes 1-3 along them).  Returns node numbers, [N, 3]
    coordinates and [NE, 4] connectivity in node numbers (from 1)."""
    xs, ys, zs = np.linspace(0, lx, nx + 1), np.linspace(0, ly, ny + 1), np.linspace(0, lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    # the cell's corners in the hexahedron numbering of ElHex8N3D
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    conn = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = [nid(i + a, j + b, k + d) for a, b, d in corners]
                for t in tets:
                    q = [c[v] for v in t]
                    v0 = xyz[q[0]]
                    if np.linalg.det(np.stack([xyz[q[1]] - v0, xyz[q[2]] - v0, xyz[q[3]] - v0])) < 0:
                        q[1], q[2] = q[2], q[1]
                    conn.append(q)
    return np.arange(1, len(xyz) + 1), xyz, np.asarray(conn) + 1


def taylor_tet(n=4, nz=16, side=6.4e-3, length=32.4e-3, speed=227.0, t_end=20e-6, material=None,
               precision="fp64", track_tensors=True, context=None, freq=10, dt_method=Explicit.PowerIteration):
    """Impact of a square copper bar (side x side x length, axis along y)
    meshed with 4-node tetrahedra (ElTet4N3D, six per cell of an n x n x nz
    grid); bottom face restrained along the axis, initial speed -speed on
    every node.  The reference's tetrahedron has no characteristic length
    (0), so the time step comes from the power iteration."""
    p = dict(COPPER)
    if material:
        p.update(material)
    model = DynELA("Taylor-Tet", context=context)
    nums, xyz, conn = _tet_box(n, nz, n, side, length, side)
    model.createNodes(nums, xyz)
    model.setDefaultElement(Element.ElTet4N3D)
    ne = conn.shape[0]
    model.createElements(np.arange(1, ne + 1), conn)
    allES = ElementSet("ES_All")
    model.add(allES, 1, ne)
    model.add(_jc_material(p), allES)
    bottom = BoundaryRestrain("BC_bottom")
    bottom.setValue(0, 1, 0)
    model.attachConstantBC(bottom, _set(nums[xyz[:, 1] == 0.0], "NS_Bottom"))
    v = BoundarySpeed("BC_speed")
    v.setValue(0, -speed, 0)
    model.attachInitialBC(v, _set(nums, "NS_All"))
    model.add(_solver(t_end, precision, track_tensors, freq, dt_method))
    return model


def necking_3d(n=16, nz=60, radius=6.413e-3, length=53.334e-3, taper=0.02, pull=1.0, t_end=50e-6,
               material=None, precision="fp64", track_tensors=True, context=None, freq=10, dt_method=None):
    """Round bar in tension: radius reduced by `taper` at mid-length with a
    cosine profile, bottom face restrained along the axis (y), top face
    pulled at constant speed `pull` (BoundarySpeed, constant)."""
    p = dict(NECKING)
    if material:
        p.update(material)
    model = DynELA("BarNecking-3D", context=context)
    ys = np.linspace(0.0, length, nz + 1)
    rad = radius * (1.0 - taper * 0.5 * (1.0 + np.cos(2.0 * np.pi * (ys / length - 0.5))))
    nums, xyz, conn, nl = _hex_cylinder(n, nz, rad, ys)
    model.createNodes(nums, xyz)
    model.setDefaultElement(Element.ElHex8N3D)
    ne = conn.shape[0]
    model.createElements(np.arange(1, ne + 1), conn)
    allES = ElementSet("ES_All")
    model.add(allES, 1, ne)
    model.add(_voce_material(p), allES)
    bottom = BoundaryRestrain("BC_bottom")
    bottom.setValue(0, 1, 0)
    model.attachConstantBC(bottom, _set(range(1, nl + 1), "NS_Bottom"))
    top = BoundarySpeed("BC_top")
    top.setValue(0, pull, 0)
    model.attachConstantBC(top, _set(range(nz * nl + 1, (nz + 1) * nl + 1), "NS_Top"))
    model.add(_solver(t_end, precision, track_tensors, freq, dt_method))
    return model


def _voce_material(p):
    law = VoceLinearLaw("Voce")
    law.setParameters(p["A"], p["Q"], p["b"], p["H"])
    mat = Material("bar")
    mat.setHardeningLaw(law)
    mat.youngModulus, mat.poissonRatio, mat.density = p["young"], p["poisson"], p["density"]
    mat.heatCapacity, mat.taylorQuinney, mat.T0 = p["heatCapacity"], p["taylorQuinney"], p["T0"]
    return mat


def necking_axi(nr=10, nz=80, radius=6.413e-3, length=53.334e-3, taper=0.02, pull=1.0, t_end=50e-6,
                material=None, precision="fp64", track_tensors=True, context=None, freq=10, dt_method=None,
                plane=False):
    """Axisymmetric round bar in tension (4-node axisymmetric quads): the
    necking_3d geometry, loads and law on the (r, y) half-plane, nodes
    restrained radially on the axis.  plane=True: the same half-plane as a
    plane-strain plate of 4-node plane quads (Samples/PlateNecking), its
    symmetry line restrained in x."""
    p = dict(NECKING)
    if material:
        p.update(material)
    model = DynELA("PlateNecking" if plane else "BarNecking-Axi", context=context)
    ys = np.linspace(0.0, length, nz + 1)
    rad = radius * (1.0 - taper * 0.5 * (1.0 + np.cos(2.0 * np.pi * (ys / length - 0.5))))
    i, j = np.meshgrid(np.arange(nr + 1), np.arange(nz + 1))
    nums = (j * (nr + 1) + i + 1).ravel()
    xyz = np.stack([(i / nr * rad[j]).ravel(), ys[j].ravel(), np.zeros(nums.size)], axis=1)
    model.createNodes(nums, xyz)
    model.setDefaultElement(Element.ElQua4N2D if plane else Element.ElQua4NAx)
    ei, ej = np.meshgrid(np.arange(nr), np.arange(nz))
    n1 = ei + ej * (nr + 1) + 1
    conn = np.stack([n1, n1 + 1, ei + (ej + 1) * (nr + 1) + 2, ei + (ej + 1) * (nr + 1) + 1], axis=-1)
    model.createElements(np.arange(1, nr * nz + 1), conn.reshape(-1, 4))
    allES = ElementSet("ES_All")
    model.add(allES, 1, nr * nz)
    model.add(_voce_material(p), allES)
    bottom = BoundaryRestrain("BC_bottom")
    bottom.setValue(0, 1, 1)
    model.attachConstantBC(bottom, _set(range(1, nr + 2), "NS_Bottom"))
    axis = BoundaryRestrain("BC_axis")
    axis.setValue(1, 0, 1)
    model.attachConstantBC(axis, _set(range(1, nums.size + 1, nr + 1), "NS_Axis"))
    top = BoundarySpeed("BC_top")
    top.setValue(0, pull, 0)
    model.attachConstantBC(top, _set(range(nz * (nr + 1) + 1, (nz + 1) * (nr + 1) + 1), "NS_Top"))
    model.add(_solver(t_end, precision, track_tensors, freq, dt_method))
    return model


def necking_plane(nx=10, ny=80, **kw):
    """Plane-strain plate in tension (ElQua4N2D): necking_axi(plane=True)."""
    return necking_axi(nr=nx, nz=ny, plane=True, **kw)


def serve_params(n_jobs, seed=20261203, lo=1000, hi=500000):
    """Seeded serving mix: kind {axisymmetric quads, 3D hexes} x {impact,
    necking} uniformly, element count log-uniform in [lo, hi], end time
    U(20, 100) us; impacts draw the sweep's copper Johnson-Cook scalings and
    speed, necking bars scale the Voce A, Q, b, H by U(0.8, 1.2)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_jobs):
        dim = "axi" if rng.random() < 0.5 else "3d"
        load = "impact" if rng.random() < 0.5 else "necking"
        ne = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        t_end = float(rng.uniform(20e-6, 100e-6))
        scale = rng.uniform(0.8, 1.2, size=5)
        speed = float(rng.uniform(150.0, 300.0))
        if load == "impact":
            mat = {k: COPPER[k] * s for k, s in zip(("A", "B", "n", "C", "m"), scale)}
            aspect = 10.0
        else:
            mat = {k: NECKING[k] * s for k, s in zip(("A", "Q", "b", "H"), scale[:4])}
            aspect = 8.0
        if dim == "axi":
            nr = max(2, int(round(np.sqrt(ne / aspect))))
            shape = (nr, max(2, int(round(ne / nr))))
        else:
            nc = max(2, int(round((ne / (1.5 * aspect)) ** (1.0 / 3.0))))
            shape = (nc, max(2, int(round(ne / (3 * nc * nc)))))
        jobs.append(dict(kind=dim, load=load, shape=shape, material=mat, speed=speed, t_end=t_end))
    return jobs


def build_serve_job(job, precision="mixed", track_tensors=False, context=None):
    a, b = job["shape"]
    kw = dict(t_end=job["t_end"], material=job["material"], precision=precision, track_tensors=track_tensors,
              context=context)
    if job["load"] == "impact":
        if job["kind"] == "axi":
            return taylor_axi(a, b, speed=job["speed"], **kw)
        return taylor_3d(a, b, speed=job["speed"], **kw)
    if job["kind"] == "axi":
        return necking_axi(a, b, **kw)
    return necking_3d(a, b, **kw)


def sweep_params(n_jobs=256, seed=20261203, lo=2000, hi=60000):
    """Seeded parameter sweep around the copper impact: kind (axi / 3D, p=0.5),
    element count log-uniform in [lo, hi], Johnson-Cook A, B, n, C, m each
    scaled by U(0.8, 1.2), impact speed U(150, 300) m/s."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_jobs):
        kind = "axi" if rng.random() < 0.5 else "3d"
        ne = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        scale = rng.uniform(0.8, 1.2, size=5)
        speed = float(rng.uniform(150.0, 300.0))
        mat = {k: COPPER[k] * s for k, s in zip(("A", "B", "n", "C", "m"), scale)}
        if kind == "axi":
            nr = max(2, int(round(np.sqrt(ne / 10.0))))
            nz = max(2, int(round(ne / nr)))
            shape = (nr, nz)
        else:
            nc = max(2, int(round((ne / 15.0) ** (1.0 / 3.0))))
            nz = max(2, int(round(ne / (3 * nc * nc))))
            shape = (nc, nz)
        jobs.append(dict(kind=kind, shape=shape, material=mat, speed=speed))
    return jobs


def build_sweep_job(job, t_end=80e-6, precision="mixed", track_tensors=False, context=None):
    if job["kind"] == "axi":
        nr, nz = job["shape"]
        return taylor_axi(nr, nz, speed=job["speed"], t_end=t_end, material=job["material"],
                          precision=precision, track_tensors=track_tensors, context=context)
    nc, nz = job["shape"]
    return taylor_3d(nc, nz, speed=job["speed"], t_end=t_end, material=job["material"], precision=precision,
                     track_tensors=track_tensors, context=context)
