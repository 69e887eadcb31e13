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

HEX, QUAD_AXI = 4, 3
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


def _assembled_operator(a, mass):
    """M^-1 (I + K) in the reference's DOF order (node * dim + d), from the
    oracle's element matrices at the initial coordinates."""
    elem_type = a.elem_type
    dim = 3 if elem_type == HEX else 2
    n = a.coords.shape[1]
    K = np.zeros((dim * n, dim * n))
    for conn in a.conn:
        Ke = element_stiffness(elem_type, a.coords[:, conn].T, a.material.young, a.material.poisson)
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
