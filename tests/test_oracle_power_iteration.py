"""Power-iteration time step of the CPU oracle (no GPU): the element
stiffness (Element3D.C:303-399, ElementPlane.C:182-292 with
Material.C:98-152) against closed forms, and the iteration
(Model.C:942-1009) against the dominant eigenvalue of the assembled
operator M^-1 (I + K)."""
import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import F_PLASTIC_STRAIN, F_TEMPERATURE, OracleJob, element_stiffness

HEX, QUAD_AXI, QUAD_PLANE = 4, 3, 1
E, NU = 200000.0, 0.3
FCT = E / ((1 + NU) * (1 - 2 * NU))


def box(a, b, c):
    return np.array([[0, 0, 0], [a, 0, 0], [a, b, 0], [0, b, 0], [0, 0, c], [a, 0, c], [a, b, c], [0, b, c]], float)


def test_hex_stiffness_symmetric_with_rigid_null_space():
    X = box(2.0, 3.0, 5.0) + np.array([0.3, -0.2, 0.1])
    X[6] += [0.2, 0.1, -0.3]  # a distorted hex
    K = element_stiffness(HEX, X, E, NU)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    for d in range(3):  # translations
        u = np.zeros(24)
        u[d::3] = 1.0
        assert np.abs(K @ u).max() <= 1e-9 * np.abs(K).max()
    for axis in range(3):  # infinitesimal rotations u = w x X
        w = np.zeros(3)
        w[axis] = 1.0
        u = np.cross(w, X).ravel()
        assert np.abs(K @ u).max() <= 1e-9 * np.abs(K).max() * np.abs(X).max()
    # positive semi-definite, six rigid modes
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() > -1e-9 * ev.max()
    assert np.sum(ev < 1e-9 * ev.max()) == 6


def test_hex_stiffness_uniaxial_stretch_and_shear_forces():
    """u = eps x on a box: sigma = C (eps, 0, ...) is uniform, so each node of
    a face carries a quarter of sigma_nn times the face area.  Simple shear
    u_x = g y: the reference's shear diagonal (1 - 2 nu) gives
    sigma_xy = E / (1 + nu) g."""
    a, b, c = 2.0, 3.0, 5.0
    X = box(a, b, c)
    K = element_stiffness(HEX, X, E, NU)
    eps = 1e-3
    u = np.zeros(24)
    u[0::3] = eps * X[:, 0]
    f = (K @ u).reshape(8, 3)
    sx, sy, sz = FCT * (1 - NU) * eps, FCT * NU * eps, FCT * NU * eps
    np.testing.assert_allclose(f[:, 0], np.where(X[:, 0] > 0, 1, -1) * sx * b * c / 4, rtol=1e-12)
    np.testing.assert_allclose(f[:, 1], np.where(X[:, 1] > 0, 1, -1) * sy * a * c / 4, rtol=1e-12)
    np.testing.assert_allclose(f[:, 2], np.where(X[:, 2] > 0, 1, -1) * sz * a * b / 4, rtol=1e-12)
    g = 2e-3
    u = np.zeros(24)
    u[0::3] = g * X[:, 1]
    f = (K @ u).reshape(8, 3)
    sxy = E / (1 + NU) * g
    np.testing.assert_allclose(f[:, 0], np.where(X[:, 1] > 0, 1, -1) * sxy * a * c / 4, rtol=1e-12)
    np.testing.assert_allclose(f[:, 1], np.where(X[:, 0] > 0, 1, -1) * sxy * b * c / 4, rtol=1e-12)
    np.testing.assert_allclose(f[:, 2], 0.0, atol=1e-9 * sxy * a * b)


def test_axi_stiffness_closed_forms():
    """Ring element [r0, r1] x [z0, z1]: an axial translation is a null
    mode; the uniform radial expansion u_r = eps r (eps_rr = eps_tt = eps)
    has sigma_rr = sigma_tt = f eps, sigma_zz = 2 nu f eps and nodal forces
    f_r = +-f eps 2 pi r_a (z1 - z0) / 2 and
    f_z = +-sigma_zz 2 pi int_{r0}^{r1} r L_a(r) dr, exact under 2 x 2 Gauss."""
    r0, r1, z0, z1 = 1.5, 2.5, 0.4, 1.9
    X = np.array([[r0, z0, 0], [r1, z0, 0], [r1, z1, 0], [r0, z1, 0]], float)
    K = element_stiffness(QUAD_AXI, X, E, NU)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    u = np.zeros(8)
    u[1::2] = 1.0
    assert np.abs(K @ u).max() <= 1e-10 * np.abs(K).max()
    eps = 1e-3
    u = np.zeros(8)
    u[0::2] = eps * X[:, 0]
    f = (K @ u).reshape(4, 2)
    fr = np.where(X[:, 0] == r1, 1, -1) * FCT * eps * 2 * np.pi * X[:, 0] * (z1 - z0) / 2
    inner = np.where(X[:, 0] == r1, (r1 - r0) * (2 * r1 + r0) / 6, (r1 - r0) * (r1 + 2 * r0) / 6)
    fz = np.where(X[:, 1] == z1, 1, -1) * 2 * NU * FCT * eps * 2 * np.pi * inner
    np.testing.assert_allclose(f[:, 0], fr, rtol=1e-12)
    np.testing.assert_allclose(f[:, 1], fz, rtol=1e-12)


def _assembled_operator(a, mass, under=False):
    """M^-1 (I + K) in the reference's DOF order (node * dim + d), from the
    oracle's element matrices at the initial coordinates (under: the
    one-point under-integrated ones)."""
    elem_type = a.elem_type
    dim = 3 if elem_type == HEX else 2
    n = a.coords.shape[1]
    K = np.zeros((dim * n, dim * n))
    for conn in a.conn:
        Ke = element_stiffness(elem_type, a.coords[:, conn].T, a.material.young, a.material.poisson, under=under)
        dofs = (np.asarray(conn)[:, None] * dim + np.arange(dim)[None, :]).ravel()
        K[np.ix_(dofs, dofs)] += Ke
    m = np.repeat(mass, dim)
    return (np.eye(dim * n) + K) / m[:, None]


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_power_iteration_converges_to_dominant_eigenvalue(kind):
    """The pre-loop time step (Explicit.C:152-154, cold start): the power
    iteration's maximum approaches the dominant eigenvalue of M^-1 (I + K)
    from below, and the step is safety * Omega_s / sqrt(fmax).  The
    reference stops once the maximum changes by less than 1e-4 in one
    iteration (Model.C:994-998); with close leading eigenvalues that leaves
    it up to a few 1e-3 below the eigenvalue (a wrong factor in K or M would
    be off by far more)."""
    if kind == "axi":
        m = samples.taylor_axi(nr=3, nz=8, dt_method=dnl.Explicit.PowerIteration)
    else:
        m = samples.taylor_3d(n=2, nz=4, dt_method=dnl.Explicit.PowerIteration)
    a = m.descriptors()
    o = OracleJob(a)
    assert o.solve(0.0) == 0  # pre-loop only
    fmax, iters = o.pi_state()
    lam = np.abs(np.linalg.eigvals(_assembled_operator(a, o.mass()))).max()
    print(f"{kind}: fmax={fmax:.6e} lambda_max={lam:.6e} iterations={iters}")
    assert iters > 1
    assert fmax <= lam * (1 + 1e-12)
    assert fmax == pytest.approx(lam, rel=1e-2)
    _, dt, _ = o.scalars()
    rb = 0.8182
    omega_s = np.sqrt((12 * (1 + rb) ** 3 * (2 - rb)) / (10 + 15 * rb - rb * rb + rb ** 3 - rb ** 4))
    assert dt == pytest.approx(0.9 * omega_s / np.sqrt(fmax), rel=1e-14)


def test_power_iteration_warm_start():
    """Later time steps start from the previous iterate and maximum
    (Model.C:966-970): with the step recomputed every increment the
    iteration stops after a few steps."""
    m = samples.taylor_axi(nr=3, nz=8, t_end=2e-6, freq=1, dt_method=dnl.Explicit.PowerIteration)
    o = OracleJob(m.descriptors())
    assert o.solve(0.0) == 0
    _, cold = o.pi_state()
    assert o.solve(2e-6) == 0
    _, warm = o.pi_state()
    print(f"cold {cold} iterations, warm {warm}")
    assert warm <= 3 < cold


def test_taylor_axi_sample_with_its_power_iteration_step():
    """Samples/Taylor/Taylor-Axi/Taylor.py as published: power-iteration
    time step recomputed every increment (Taylor.py:117-118).  Published
    results (Documentation/Samples-Impacts.tex:106-107): final length
    26.52 mm, impact-face diameter 11.15 mm, centre element plastic strain
    1.78 and temperature 582.34 C."""
    steel = dict(young=206000.0, poisson=0.3, density=7.83e-09, heatCapacity=4.6e+08, taylorQuinney=0.9,
                 A=806.0, B=614.0, C=0.0089, n=0.168, m=1.1, depsp0=1.0, Tm=1540.0, T0=20.0)
    m = samples.taylor_axi(nr=5, nz=50, radius=3.2, length=32.4, speed=287000.0, t_end=80e-6, material=steel,
                           freq=1, dt_method=dnl.Explicit.PowerIteration)
    o = OracleJob(m.descriptors(), threads=4)
    assert o.solve(80e-6) == 0
    c, _, _ = o.nodes()
    L_f = c[1].max()
    D_f = 2 * c[0][:6].max()
    ep = o.ip(F_PLASTIC_STRAIN)[0].mean()
    T = o.ip(F_TEMPERATURE)[0].mean()
    _, _, inc = o.scalars()
    print(f"L_f={L_f:.3f} mm D_f={D_f:.3f} mm ep={ep:.3f} T={T:.1f} C increments={inc}")
    assert L_f == pytest.approx(26.52, rel=1e-3)
    assert D_f == pytest.approx(11.15, rel=1e-3)
    assert ep == pytest.approx(1.78, rel=1e-2)
    assert T == pytest.approx(582.34, rel=2e-3)


# ---- PowerIterationUnder: the one-point under-integrated stiffness ---------
# (Model.C:942-1009 with underIntegration = true; Element3D.C:118-167,
# ElementPlane.C:104-160, the centre point of ElHex8N3D.C:233-252 and
# ElQua4NAx.C:112-124.)  Pinned against the strain at the element centre
# formed by central differences of the element's multilinear displacement
# interpolant (exact for a multilinear function), independently of the
# oracle's shape tables: u^T K u = V_w eps_c^T C eps_c, with V_w the
# centre rule's weighted volume (8 det J for a box = its volume; 2 pi r_c
# times the area for an axisymmetric rectangle).

def _hooke(dim_kind):
    """Material.C:98-152 with the reference's (1 - 2 nu) shear diagonal:
    rows (xx, yy, zz, xy, yz, xz) in 3D, (rr, zz, rz, tt) axisymmetric."""
    if dim_kind == "3d":
        C = np.zeros((6, 6))
        C[:3, :3] = NU
        np.fill_diagonal(C[:3, :3], 1 - NU)
        C[3, 3] = C[4, 4] = C[5, 5] = 1 - 2 * NU
    else:
        C = np.array([[1 - NU, NU, 0, NU], [NU, 1 - NU, 0, NU], [0, 0, 1 - 2 * NU, 0], [NU, NU, 0, 1 - NU]])
    return FCT * C


def _interp(lo, hi, nodal, pt):
    """Multilinear interpolation over the box [lo, hi] of nodal values given
    at the box corners in the reference's node numbering."""
    dim = len(lo)
    signs = ([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]]
             if dim == 3 else [[-1, -1], [1, -1], [1, 1], [-1, 1]])
    t = (2 * (np.asarray(pt) - lo) / (np.asarray(hi) - lo)) - 1
    w = [np.prod([(1 + s[d] * t[d]) / 2 for d in range(dim)]) for s in signs]
    return np.tensordot(w, nodal, axes=1)


def _centre_strain(lo, hi, u, axi):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c = 0.5 * (lo + hi)
    dim = len(lo)
    grad = np.zeros((dim, dim))  # grad[i, j] = du_i / dx_j
    for j in range(dim):
        h = 1e-3 * (hi[j] - lo[j])
        e = np.zeros(dim)
        e[j] = h
        grad[:, j] = (_interp(lo, hi, u, c + e) - _interp(lo, hi, u, c - e)) / (2 * h)
    if dim == 3:
        return np.array([grad[0, 0], grad[1, 1], grad[2, 2], grad[0, 1] + grad[1, 0], grad[1, 2] + grad[2, 1],
                         grad[0, 2] + grad[2, 0]])
    uc = _interp(lo, hi, u, c)
    return np.array([grad[0, 0], grad[1, 1], grad[0, 1] + grad[1, 0], uc[0] / c[0] if axi else 0.0])


def test_under_stiffness_hex_energy_rank_and_hourglass():
    """Box hex: u^T K u = V eps_c^T C eps_c for random u; K symmetric, PSD
    of rank 6 (one point, six strain components: six rigid modes and
    twelve hourglass modes in its null space), the hourglass vectors
    (products of two or three local signs, per direction) among them while
    the fully integrated matrix resists them; on linear fields the two
    matrices agree (the centre rule is exact for constant strain on a
    parallelepiped)."""
    rng = np.random.default_rng(3)
    lo, hi = np.array([0.3, -0.2, 0.1]), np.array([2.3, 2.8, 5.1])
    X = box(*(hi - lo)) + lo
    K = element_stiffness(HEX, X, E, NU, under=True)
    Kf = element_stiffness(HEX, X, E, NU)
    V = np.prod(hi - lo)
    C = _hooke("3d")
    for _ in range(5):
        u = rng.standard_normal((8, 3))
        eps = _centre_strain(lo, hi, u, False)
        assert u.ravel() @ K @ u.ravel() == pytest.approx(V * eps @ C @ eps, rel=1e-8)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() > -1e-9 * ev.max()
    assert np.sum(ev > 1e-9 * ev.max()) == 6
    s = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    for h in (s[:, 0] * s[:, 1], s[:, 1] * s[:, 2], s[:, 0] * s[:, 2], s[:, 0] * s[:, 1] * s[:, 2]):
        for d in range(3):
            u = np.zeros((8, 3))
            u[:, d] = h
            assert np.abs(K @ u.ravel()).max() <= 1e-10 * np.abs(K).max()
            assert np.abs(Kf @ u.ravel()).max() > 1e-3 * np.abs(Kf).max()
    # a distorted parallelepiped (affine map of the box) and a linear field
    A = np.array([[1.0, 0.2, -0.1], [0.1, 0.9, 0.3], [-0.2, 0.1, 1.1]])
    Xp = X @ A.T
    G = rng.standard_normal((3, 3))
    u = (Xp @ G.T).ravel()
    Ku = element_stiffness(HEX, Xp, E, NU, under=True) @ u
    np.testing.assert_allclose(Ku, element_stiffness(HEX, Xp, E, NU) @ u, rtol=0, atol=1e-9 * np.abs(Ku).max())


@pytest.mark.parametrize("axi", [True, False])
def test_under_stiffness_quad_energy_and_rank(axi):
    """Axisymmetric ring / plane-strain rectangle: u^T K u = V_w eps_c^T C
    eps_c (axisymmetric: V_w = 2 pi r_c x area, the hoop strain u_r / r at
    the centre; plane strain: the area, unit thickness); rank 4 / 3; an
    axial translation (and for plane strain a rigid motion) is a null
    mode."""
    rng = np.random.default_rng(5)
    lo, hi = np.array([1.5, 0.4]), np.array([2.5, 1.9])
    X = np.array([[lo[0], lo[1], 0], [hi[0], lo[1], 0], [hi[0], hi[1], 0], [lo[0], hi[1], 0]])
    K = element_stiffness(QUAD_AXI if axi else QUAD_PLANE, X, E, NU, under=True)
    area = np.prod(hi - lo)
    Vw = 2 * np.pi * 0.5 * (lo[0] + hi[0]) * area if axi else area
    C = _hooke("axi")
    if not axi:
        C = C[:3, :3]
    for _ in range(5):
        u = rng.standard_normal((4, 2))
        eps = _centre_strain(lo, hi, u, axi)
        if not axi:
            eps = eps[:3]
        assert u.ravel() @ K @ u.ravel() == pytest.approx(Vw * eps @ C @ eps, rel=1e-8)
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() > -1e-9 * ev.max()
    assert np.sum(ev > 1e-9 * ev.max()) == (4 if axi else 3)
    u = np.zeros(8)
    u[1::2] = 1.0
    assert np.abs(K @ u).max() <= 1e-10 * np.abs(K).max()
    if not axi:
        u = np.cross([0, 0, 1.0], X)[:, :2].ravel()
        assert np.abs(K @ u).max() <= 1e-10 * np.abs(K).max() * 3


@pytest.mark.parametrize("kind", ["axi", "3d"])
def test_power_iteration_under_converges_to_dominant_eigenvalue(kind):
    """PowerIterationUnder's pre-loop step: the maximum approaches the
    dominant eigenvalue of M^-1 (I + K_under) from below, below the fully
    integrated operator's (the hourglass stiffness is left out), and the
    step is safety * Omega_s / sqrt(fmax).  The under-integrated operator's
    leading eigenvalues lie closer together, so the reference's test (the
    maximum changed by less than 1e-4 in one iteration) stops the cold start
    further below the eigenvalue (2.7 % on the axisymmetric mesh); warm
    starts over the next increments close the gap."""
    mk = (lambda **k: samples.taylor_axi(nr=3, nz=8, **k)) if kind == "axi" else \
        (lambda **k: samples.taylor_3d(n=2, nz=4, **k))
    m = mk(dt_method=dnl.Explicit.PowerIterationUnder)
    a = m.descriptors()
    o = OracleJob(a)
    assert o.solve(0.0) == 0
    fmax, iters = o.pi_state()
    lam = np.abs(np.linalg.eigvals(_assembled_operator(a, o.mass(), under=True))).max()
    lam_full = np.abs(np.linalg.eigvals(_assembled_operator(a, o.mass()))).max()
    print(f"{kind}: fmax={fmax:.6e} lambda_under={lam:.6e} lambda_full={lam_full:.6e} iterations={iters}")
    assert iters > 1
    assert fmax <= lam * (1 + 1e-12)
    assert fmax == pytest.approx(lam, rel=5e-2)
    assert lam < lam_full
    _, dt, _ = o.scalars()
    rb = 0.8182
    omega_s = np.sqrt((12 * (1 + rb) ** 3 * (2 - rb)) / (10 + 15 * rb - rb * rb + rb ** 3 - rb ** 4))
    assert dt == pytest.approx(0.9 * omega_s / np.sqrt(fmax), rel=1e-14)
