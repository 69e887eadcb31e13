"""The 4-node tetrahedron (ElTet4N3D.C) in the CPU oracle (no GPU).

The reference's ElTet4N3D has one integration point (the centroid, weight
1/6, linear shape functions) and runs only under a power-iteration time
step: its characteristic length is 0 (ElTet4N3D.h:45), so its Courant step
is 0.  Pins, independent of the oracle's tables: the element stiffness of a
constant-strain tetrahedron from the affine displacement field of its four
nodes (u^T K u = V eps^T C eps), its rank and rigid null space; the lumped
masses rho V / 4; free flight; the reference's node-field extrapolation
quirk (ElTet4N3D.C:33-60: the first entry of each node's four-entry vector);
and a tetrahedral impact that yields, with energy accounted for."""
import numpy as np
import pytest

import dnlPython as dnl
from dnlPython import samples
from oracle.bridge import F_PLASTIC_STRAIN, OracleJob, element_stiffness, extrapolation

TET = 5
E, NU = 200000.0, 0.3
FCT = E / ((1 + NU) * (1 - 2 * NU))


def _hooke3():
    C = np.zeros((6, 6))
    C[:3, :3] = NU
    np.fill_diagonal(C[:3, :3], 1 - NU)
    C[3, 3] = C[4, 4] = C[5, 5] = 1 - 2 * NU  # the reference's shear diagonal (Material.C:98-152)
    return FCT * C


def _affine_strain(X, u):
    """Engineering strain (xx, yy, zz, xy, yz, xz) of the affine field through
    the four nodal displacements."""
    A = (X[1:] - X[0]).T  # columns: edge vectors
    G = (u[1:] - u[0]).T @ np.linalg.inv(A)  # G[i, j] = du_i / dx_j
    return np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])


def test_tet_stiffness_energy_rank_and_rigid_modes():
    rng = np.random.default_rng(7)
    X = np.array([[0.1, 0.2, -0.1], [2.1, 0.3, 0.2], [0.4, 1.7, 0.1], [0.2, 0.5, 1.9]])
    V = np.linalg.det((X[1:] - X[0]).T) / 6
    assert V > 0
    K = element_stiffness(TET, X, E, NU)
    C = _hooke3()
    for _ in range(5):
        u = rng.standard_normal((4, 3))
        eps = _affine_strain(X, u)
        assert u.ravel() @ K @ u.ravel() == pytest.approx(V * eps @ C @ eps, rel=1e-10)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert ev.min() > -1e-9 * ev.max()
    assert np.sum(ev > 1e-9 * ev.max()) == 6
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1.0
        assert np.abs(K @ np.cross(w, X).ravel()).max() <= 1e-10 * np.abs(K).max() * 3
    # the under-integration point of the tetrahedron is its integration point
    np.testing.assert_array_equal(element_stiffness(TET, X, E, NU, under=True), K)


def test_tet_extrapolation_is_the_references_first_entry():
    """ElTet4N3D.C:33-60: node a's vector starts with (1 + 3 sqrt5)/4 for
    a = 0 and (1 - sqrt5)/4 otherwise; with one integration point only that
    entry is used (Node.C:250-260)."""
    Et = extrapolation(TET)
    assert Et.shape == (4, 1)
    np.testing.assert_array_equal(Et[:, 0], [(1 + 3 * np.sqrt(5)) / 4] + [(1 - np.sqrt(5)) / 4] * 3)


def test_tet_masses_free_flight_and_courant_refused():
    # no restraint: free flight of the whole bar
    m2 = samples.taylor_tet(n=2, nz=4, speed=100.0, t_end=2e-6)
    m2._constant.clear()
    m2._invalidate()
    o = OracleJob(m2.descriptors())
    assert o.solve(2e-6) == 0
    mass = o.mass()
    V = 6.4e-3 * 6.4e-3 * 32.4e-3
    assert mass.sum() == pytest.approx(8930.0 * V, rel=1e-13)
    c, d, v = o.nodes()
    t, _, _ = o.scalars()
    np.testing.assert_allclose(d[1], -100.0 * t, rtol=1e-12)
    np.testing.assert_allclose(d[[0, 2]], 0.0, atol=1e-15)
    assert np.abs(o.ip(F_PLASTIC_STRAIN)).max() == 0.0
    # the Courant step of a tetrahedral model is refused (its length is 0)
    m3 = samples.taylor_tet(n=2, nz=4, t_end=2e-6, dt_method=dnl.Explicit.Courant)
    with pytest.raises(RuntimeError):
        OracleJob(m3.descriptors())


def test_tet_element_masses_are_a_quarter_of_rho_v():
    """Element::computeMassMatrix (Element.C:304-339) over the one point:
    each node takes rho N_a w detJ = rho V / 4."""
    m = samples.taylor_tet(n=1, nz=1, t_end=1e-9)
    a = m.descriptors()
    o = OracleJob(a)
    assert o.solve(0.0) == 0
    X = a.coords.T
    vol = [np.linalg.det((X[c[1:]] - X[c[0]]).T) / 6 for c in a.conn]
    expect = np.zeros(len(X))
    for c, vv in zip(a.conn, vol):
        expect[c] += 8930.0 * vv / 4
    np.testing.assert_allclose(o.mass(), expect, rtol=1e-13)


def test_tet_impact_yields_and_conserves_energy():
    """The tetrahedral impact under its power-iteration step: the bar yields
    at the impact face, and KE + IE falls short of the initial kinetic
    energy only by the Chung-Hulbert dissipation of the coarse mesh's
    highest modes (the hex mesh of the same cell size loses the same 4.3 %
    in 10 us), which halves when the mesh is refined twice."""
    bal = []
    for n, nz in ((3, 12), (4, 24)):
        m = samples.taylor_tet(n=n, nz=nz, t_end=10e-6)
        o = OracleJob(m.descriptors())
        assert o.solve(10e-6) == 0
        ep = o.ip(F_PLASTIC_STRAIN)
        assert ep.max() > 0.1
        ke, ie = o.energies()[:2]
        ke0 = 0.5 * o.mass().sum() * 227.0 ** 2
        bal.append((ke0 - ke - ie) / ke0)
        print(f"{n}x{n}x{nz}: ep max {ep.max():.3f}, KE {ke:.4e}, IE {ie:.4e}, KE0 {ke0:.4e}, lost {bal[-1]:.3e}")
    assert 0 < bal[0] < 5e-2
    assert 0 < bal[1] < 0.6 * bal[0]
