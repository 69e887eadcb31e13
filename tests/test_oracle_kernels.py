"""Pins the CPU oracle function by function, without the GPU.

Each check is something the mathematics or the reference itself fixes:
  - the symmetric eigensolver against the reference's own dsyevd3 built from
    Sources/dnlMaths/Eigen3x3.C (bit-for-bit), and the eigen-identities;
  - the polar decomposition / log strain against scipy's polar and logm,
    plus invariants (R orthogonal, F = R exp(dE), pure stretch, rotation);
  - shape tables against the trilinear / bilinear shape functions;
  - Johnson-Cook slope against the reference's finite-difference recipe
    (JohnsonCookLaw.C:342-358) and special branches;
  - radial return against the closed form of linear hardening and against a
    brute-force root of the scalar return equation;
  - Chung-Hulbert constants against the method's defining property: a
    triple root -rho_b of the amplification matrix at bifurcation;
  - characteristic lengths on boxes, rectangles and parallelepipeds.
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.linalg as sla
from scipy.optimize import brentq

from oracle import bridge

lib = bridge.load()
P = bridge.p
_dp = ctypes.POINTER(ctypes.c_double)


def eig_oracle(A):
    Q, w = np.zeros(9), np.zeros(3)
    lib.oracle_eig3(P(np.ascontiguousarray(A, dtype=np.float64).ravel()), P(Q), P(w))
    return Q.reshape(3, 3), w


def polar(F):
    dE, R = np.zeros(6), np.zeros(9)
    lib.oracle_polar_lnU(P(np.ascontiguousarray(F, dtype=np.float64).ravel()), P(dE), P(R))
    return dE, R.reshape(3, 3)


def sym(v):
    return np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])


def random_matrices(n, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        m = k % 6
        if m == 0:
            A = rng.normal(size=(3, 3))
            A = A + A.T
        elif m == 1:
            F = np.eye(3) + rng.normal(scale=10 ** rng.uniform(-8, -1), size=(3, 3))
            A = F.T @ F
        elif m == 2:
            F = np.diag(1 + rng.normal(scale=1e-3, size=3))
            A = F.T @ F
        elif m == 3:
            A = np.diag(rng.normal(size=3))
        elif m == 4:  # axisymmetric: F(2,2) decoupled
            F = np.eye(3)
            F[:2, :2] += rng.normal(scale=1e-3, size=(2, 2))
            F[2, 2] += rng.normal(scale=1e-4)
            A = F.T @ F
        else:  # repeated eigenvalues
            Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
            lam = rng.uniform(0.5, 2.0)
            A = Q @ np.diag([lam, lam, rng.uniform(0.5, 2.0)]) @ Q.T
        out.append(A)
    return out


@pytest.mark.skipif(not os.path.exists(bridge.REF_EIGEN), reason="reference eigensolver not built (needs the reference tree)")
def test_eig3_bitwise_equals_reference_dsyevd3():
    ref = ctypes.CDLL(bridge.REF_EIGEN)
    ref.ref_dsyevd3.argtypes = [_dp, _dp, _dp]
    for A in random_matrices(30000):
        a = np.ascontiguousarray(A, dtype=np.float64).ravel()
        q1, w1 = eig_oracle(a)
        q2, w2 = np.zeros(9), np.zeros(3)
        ref.ref_dsyevd3(P(a), P(q2), P(w2))
        assert np.array_equal(q1.ravel().view(np.int64), q2.view(np.int64))
        assert np.array_equal(w1.view(np.int64), w2.view(np.int64))


def test_eig3_identities():
    """A Q = Q diag(w) and the eigenvalues for every family; orthonormal
    eigenvectors whenever the spectrum is simple.  With an exactly repeated
    eigenvalue the Cuppen secular step of Eigen3x3.C:454-634 may return
    non-orthogonal vectors inside the degenerate eigenspace; the oracle keeps
    that behaviour (it is bit-equal to the reference above)."""
    for k, A in enumerate(random_matrices(3000, seed=11)):
        Q, w = eig_oracle(A)
        scale = max(1.0, np.abs(A).max())
        np.testing.assert_allclose(A @ Q, Q * w, atol=2e-13 * scale)
        np.testing.assert_allclose(np.sort(w), np.linalg.eigvalsh(A), atol=2e-13 * scale)
        ev = np.sort(np.linalg.eigvalsh(A))
        gap = np.min(np.diff(ev)) / max(np.abs(ev).max(), 1e-300)
        if k % 6 != 5 and gap > 1e-8:
            # orthogonality error of eigenvectors grows like eps / relative gap
            np.testing.assert_allclose(Q.T @ Q, np.eye(3), atol=max(2e-14, 4e-16 / gap))


def test_polar_against_scipy():
    rng = np.random.default_rng(3)
    cases = [np.array([[1, 5, 3], [7, 4, 5], [6, 8, -1]], dtype=float)]  # polarDecompose.C:7 input
    for k in range(400):
        scale = 10 ** rng.uniform(-6, -0.5)
        cases.append(np.eye(3) + rng.normal(scale=scale, size=(3, 3)))
    for F in cases:
        dE, R = polar(F)
        Rs, Us = sla.polar(F)  # F = Rs Us
        logU = sla.logm(Us).real
        tol = 1e-12 * max(1.0, np.abs(logU).max())
        np.testing.assert_allclose(R, Rs, atol=1e-12)
        np.testing.assert_allclose(sym(dE), logU, atol=tol)


# Deformation gradients with a nearly repeated pair of principal stretches
# (equal lateral stretches of a uniaxial wave, tiny shears) on which the
# Kopp / Cuppen solver returns NaN eigenvectors; found by brute force over
# near-identity F with random sparsity and exponents.
DEGENERATE_F = [
    [0.9999977554099807, -6.0866254721819887e-17, -2.5673284452683152e-11, -3.8754985562266079e-18,
     0.99928435481611444, -2.0196019442639356e-21, 5.3641098223804799e-20, 2.0071721319477389e-10,
     0.9999977554099807],
    [1, -5.197663647861295e-15, 0, 1.5588983398066094e-08, 1.0000000000000004, -9.724847248502339e-14, 0, 0,
     1.0000144463341012],
    [1.0000000000000002, 0, 1.6394305225272209e-13, 0, 0.99999375465211049, -2.6574145181028933e-13,
     3.0252562602304473e-12, 0, 1.0000000000000002],
    [0.99999999998633027, 0, -1.8783223923328994e-16, -1.2823616560458548e-09, 0.99999999998633027, 0, 0,
     -6.7279513642765257e-13, 0.99992631291167289],
    [0.9999999996621316, 0, -6.1559412697832952e-08, 0, 0.99993721886807174, 1.3501324046147427e-12,
     -2.9045539929949685e-12, 0, 0.9999999996621316],
]


# Near-uniaxial F on which the Kopp / Cuppen solver returns finite but
# non-orthonormal eigenvectors (orthogonality error 1e-6 .. 8e-2).
NONORTHO_F = [
    [0.9999979082640562, 0.0, 0.0, -5.150127289795424e-14, 1.000006275207761, 0.0, 0.0, 7.600647450335941e-14,
     0.9999979082640562],
    [0.9999998171516018, 0.0, -5.793171832096426e-15, 1.3557360005053231e-15, 1.00000054854519,
     2.482959666128738e-15, 0.0, 0.0, 0.999999817151603],
    [0.9996046643630397, -4.0170422591200285e-09, 0.0, 0.0, 1.0011860025955537, 2.9862197451576665e-09, 0.0, 0.0,
     0.9996046658014821],
    [0.999999998796853, 1.875980360862417e-15, 2.3646520162924696e-15, 3.267901182264114e-15,
     1.0000000036094319, 3.37880430528116e-15, 0.0, 0.0, 0.9999999987968561],
]


def _ftf(F):
    """F^T F with the summation order of Tensor2::buildFTF (no fused ops)."""
    C = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            C[i, j] = F[0, i] * F[0, j] + F[1, i] * F[1, j] + F[2, i] * F[2, j]
    return C


@pytest.mark.skipif(not os.path.exists(bridge.REF_EIGEN), reason="reference eigensolver not built (needs the reference tree)")
def test_reference_dsyevd3_fails_on_nearly_repeated_pairs():
    """The reference's own dsyevd3 (Eigen3x3.C) returns non-finite
    eigenvectors on these inputs, bit-equal to the oracle's restatement."""
    ref = ctypes.CDLL(bridge.REF_EIGEN)
    ref.ref_dsyevd3.argtypes = [_dp, _dp, _dp]
    for F in DEGENERATE_F:
        a = np.ascontiguousarray(_ftf(np.array(F).reshape(3, 3))).ravel()
        q1, w1 = eig_oracle(a)
        q2, w2 = np.zeros(9), np.zeros(3)
        ref.ref_dsyevd3(P(a), P(q2), P(w2))
        assert not np.isfinite(q2).all()
        assert np.array_equal(q1.ravel().view(np.int64), q2.view(np.int64))


@pytest.mark.skipif(not os.path.exists(bridge.REF_EIGEN), reason="reference eigensolver not built (needs the reference tree)")
def test_reference_dsyevd3_loses_orthogonality_on_near_uniaxial_stretch():
    """The reference's dsyevd3 returns finite eigenvectors that are far from
    orthonormal on these inputs, bit-equal to the oracle's restatement."""
    ref = ctypes.CDLL(bridge.REF_EIGEN)
    ref.ref_dsyevd3.argtypes = [_dp, _dp, _dp]
    for F in NONORTHO_F:
        a = np.ascontiguousarray(_ftf(np.array(F).reshape(3, 3))).ravel()
        q1, w1 = eig_oracle(a)
        q2, w2 = np.zeros(9), np.zeros(3)
        ref.ref_dsyevd3(P(a), P(q2), P(w2))
        assert np.array_equal(q1.ravel().view(np.int64), q2.view(np.int64))
        Q = q2.reshape(3, 3)
        assert np.isfinite(Q).all() and np.abs(Q.T @ Q - np.eye(3)).max() > 1e-7


def jacobi(A):
    Q, w = np.zeros(9), np.zeros(3)
    lib.oracle_jacobi_eig3(P(np.ascontiguousarray(A, dtype=np.float64).ravel()), P(Q), P(w))
    return Q.reshape(3, 3), w


def test_jacobi_eig3():
    """Cyclic Jacobi: A Q = Q diag(w), Q orthonormal, for random, repeated
    and diagonal spectra (including the degenerate F^T F above)."""
    mats = list(random_matrices(2000, seed=13))
    mats += [_ftf(np.array(F).reshape(3, 3)) for F in DEGENERATE_F + NONORTHO_F]
    mats += [np.eye(3), np.diag([2.0, 2.0, 1.0]), np.zeros((3, 3))]
    for A in mats:
        A = np.asarray(A, dtype=float)
        Q, w = jacobi(A)
        scale = max(1.0, np.abs(A).max())
        np.testing.assert_allclose(A @ Q, Q * w, atol=4e-15 * scale)
        np.testing.assert_allclose(Q.T @ Q, np.eye(3), atol=4e-15)
        np.testing.assert_allclose(np.sort(w), np.linalg.eigvalsh(A), atol=4e-15 * scale)


def test_polar_falls_back_on_nearly_repeated_pairs():
    """Where the Kopp solver fails (non-finite or non-orthonormal vectors),
    the oracle's polar decomposition (and the engine's) still returns R and
    ln U, equal to scipy's polar / logm."""
    for F in DEGENERATE_F + NONORTHO_F:
        F = np.array(F).reshape(3, 3)
        dE, R = polar(F)
        assert np.isfinite(dE).all() and np.isfinite(R).all()
        Rs, Us = sla.polar(F)
        np.testing.assert_allclose(R, Rs, atol=1e-13)
        np.testing.assert_allclose(sym(dE), sla.logm(Us).real, atol=1e-14)


def test_polar_invariants():
    rng = np.random.default_rng(5)
    for k in range(300):
        F = np.eye(3) + rng.normal(scale=0.05, size=(3, 3))
        dE, R = polar(F)
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-14)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_allclose(R @ sla.expm(sym(dE)), F, atol=1e-13)


def test_polar_pure_stretch_and_rotation():
    lam = np.array([1.001, 0.9995, 1.0002])
    dE, R = polar(np.diag(lam))
    np.testing.assert_allclose(sym(dE), np.diag(np.log(lam)), atol=1e-16)
    np.testing.assert_allclose(R, np.eye(3), atol=1e-16)
    th = 0.3
    Rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    dE, R = polar(Rot)
    assert np.abs(dE).max() < 1e-15
    np.testing.assert_allclose(R, Rot, atol=1e-15)
    dE, R = polar(np.eye(3))
    assert np.all(dE == 0.0)
    np.testing.assert_array_equal(R, np.eye(3))


def tables(elem_type):
    N, D, w = np.zeros(64), np.zeros(192), np.zeros(8)
    nip = lib.oracle_tables(elem_type, P(N), P(D), P(w))
    nn = nip
    return N[: nip * nn].reshape(nip, nn), D[: nip * nn * 3].reshape(nip, nn, 3), w[:nip]


def test_shape_tables_hex():
    N, D, w = tables(4)
    g = 1 / np.sqrt(3)
    nodes = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    for p in range(8):
        xi = nodes[p] * g  # Gauss points in the node order (ElHex8N3D.C:93-231)
        for a in range(8):
            f = (1 + nodes[a] * xi)
            assert N[p, a] == pytest.approx(f.prod() / 8, abs=1e-16)
            for c in range(3):
                o = [k for k in range(3) if k != c]
                assert D[p, a, c] == pytest.approx(nodes[a, c] * f[o].prod() / 8, abs=1e-16)
        assert N[p].sum() == pytest.approx(1.0, abs=1e-15)
        np.testing.assert_allclose(D[p].sum(axis=0), 0.0, atol=1e-16)
    np.testing.assert_array_equal(w, 1.0)


def test_shape_tables_quad():
    N, D, w = tables(3)
    g = 1 / np.sqrt(3)
    nodes = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]])
    for p in range(4):
        xi = nodes[p] * g
        for a in range(4):
            f = (1 + nodes[a] * xi)
            assert N[p, a] == pytest.approx(f.prod() / 4, abs=1e-16)
            assert D[p, a, 0] == pytest.approx(nodes[a, 0] * f[1] / 4, abs=1e-16)
            assert D[p, a, 1] == pytest.approx(nodes[a, 1] * f[0] / 4, abs=1e-16)
        assert N[p].sum() == pytest.approx(1.0, abs=1e-15)


def material(law, p, E=124e9, nu=0.34, rho=8930.0, cp=383.0, eta=0.9, T0=20.0):
    from dnlPython._lib import MaterialDesc
    m = MaterialDesc()
    m.young, m.poisson, m.density, m.heat_capacity, m.taylor_quinney, m.T0 = E, nu, rho, cp, eta, T0
    m.law = law
    for k, v in enumerate(p):
        m.p[k] = v
    return m


JC = [90e6, 292e6, 0.025, 0.31, 1.09, 1.0, 1083.0, 20.0]


def yld(m, e, de, T):
    return lib.oracle_law_yield(ctypes.addressof(m), e, de, T)


def dyld(m, e, de, T, dt):
    return lib.oracle_law_dyield(ctypes.addressof(m), e, de, T, dt)


def test_jc_slope_matches_reference_finite_difference_recipe():
    """JohnsonCookLaw::getNumericalDerivateYieldStress (JohnsonCookLaw.C:342-358):
    h = dsy/dep + dsy/ddep / dt + eta/(rho Cp) * sy * dsy/dT, here with
    central differences."""
    m = material(0, JC)
    hf = m.taylor_quinney / (m.density * m.heat_capacity)
    rng = np.random.default_rng(1)
    for _ in range(200):
        e = 10 ** rng.uniform(-5, 0.3)
        de = 10 ** rng.uniform(-2, 6)
        T = rng.uniform(0.0, 1000.0)
        dt = 10 ** rng.uniform(-9, -7)
        s = yld(m, e, de, T)
        he, hd, hT = e * 1e-6, de * 1e-6, 1e-4
        d_e = (yld(m, e + he, de, T) - yld(m, e - he, de, T)) / (2 * he)
        d_d = (yld(m, e, de + hd, T) - yld(m, e, de - hd, T)) / (2 * hd) if de > 1.0 + hd else 0.0
        d_T = (yld(m, e, de, T + hT) - yld(m, e, de, T - hT)) / (2 * hT) if 20.0 + hT < T < 1083.0 - hT else 0.0
        num = d_e + d_d / dt + hf * s * d_T
        assert dyld(m, e, de, T, dt) == pytest.approx(num, rel=2e-5, abs=1e-3 * s)


def test_jc_branches():
    m = material(0, JC)
    base = yld(m, 0.1, 1.0, 20.0)
    assert base == pytest.approx(90e6 + 292e6 * 0.1 ** 0.31, rel=1e-15)
    assert yld(m, 0.1, 0.5, 20.0) == base            # rate below depsp0: no rate term
    assert yld(m, 0.1, 1.0, 10.0) == base            # T below T0: no softening
    assert yld(m, 0.1, 1.0, 1083.0) == 0.0           # T at / above Tm: zero
    assert yld(m, 0.1, 1.0, 2000.0) == 0.0
    assert yld(m, 0.1, np.e, 20.0) == pytest.approx(base * 1.025, rel=1e-15)


def radial_return(m, dE, pressure, dt, state):
    st = np.ascontiguousarray(state, dtype=np.float64).copy()
    rc = lib.oracle_radial_return(ctypes.addressof(m), P(np.ascontiguousarray(dE, dtype=np.float64)),
                                  ctypes.c_double(pressure), ctypes.c_double(dt), P(st))
    assert rc == 0
    return st


def test_radial_return_linear_hardening_closed_form():
    """ElastoplasticLaw with n = 1: sy = A + B ep, so the return equation
    f(g) = S_tr - sqrt(3/2) 2G g - sy(ep0 + sqrt(2/3) g) is linear in g."""
    A, B = 200e6, 1e9
    m = material(1, [A, B, 1.0])
    G2 = m.young / (1 + m.poisson)
    ep0 = 0.05
    state = np.zeros(14)
    state[6], state[8], state[9], state[11] = ep0, A + B * ep0, 1e-4, 20.0
    dE = np.array([2e-3, 1e-3, 0.0, -1e-3, 0.0, -1e-3])
    dev = dE.copy()
    tr = (dE[0] + dE[3] + dE[5]) / 3
    dev[0] -= tr; dev[3] -= tr; dev[5] -= tr
    s_tr = G2 * dev
    S = np.sqrt(s_tr[0] ** 2 + s_tr[3] ** 2 + s_tr[5] ** 2 + 2 * (s_tr[1] ** 2 + s_tr[2] ** 2 + s_tr[4] ** 2))
    S_tr = np.sqrt(1.5) * S
    gamma = (S_tr - (A + B * ep0)) / (np.sqrt(1.5) * G2 + np.sqrt(2 / 3) * B)
    out = radial_return(m, dE, 0.0, 1e-8, state)
    assert out[9] == pytest.approx(gamma, rel=1e-12)
    assert out[6] == pytest.approx(ep0 + np.sqrt(2 / 3) * gamma, rel=1e-14)
    s_new = out[:6]
    Sn = np.sqrt(s_new[0] ** 2 + s_new[3] ** 2 + s_new[5] ** 2 + 2 * (s_new[1] ** 2 + s_new[2] ** 2 + s_new[4] ** 2))
    assert np.sqrt(1.5) * Sn == pytest.approx(A + B * out[6], rel=1e-12)   # back on the yield surface


def test_radial_return_johnson_cook_brute_force_root():
    m = material(0, JC)
    G2 = m.young / (1 + m.poisson)
    hf = m.taylor_quinney / (m.density * m.heat_capacity)
    rng = np.random.default_rng(9)
    dt = 3e-8
    for _ in range(40):
        ep0 = rng.uniform(1e-3, 1.0)
        T0 = rng.uniform(20.0, 600.0)
        s0 = np.zeros(6)
        sy0 = yld(m, ep0, 1e3, T0)
        state = np.zeros(14)
        state[6], state[7], state[8], state[9], state[11] = ep0, 1e3, sy0, rng.uniform(1e-6, 1e-4), T0
        dE = rng.normal(scale=2e-3, size=6)
        out = radial_return(m, dE, 0.0, dt, state)
        dev = dE.copy()
        tr = (dE[0] + dE[3] + dE[5]) / 3
        dev[0] -= tr; dev[3] -= tr; dev[5] -= tr
        s_tr = s0 + G2 * dev
        S = np.sqrt(s_tr[0] ** 2 + s_tr[3] ** 2 + s_tr[5] ** 2 + 2 * (s_tr[1] ** 2 + s_tr[2] ** 2 + s_tr[4] ** 2))
        S_tr = np.sqrt(1.5) * S
        if S_tr <= sy0:
            assert out[9] == state[9]  # elastic: gamma untouched
            continue

        def f(g):
            # temperature at the converged point: T0 + g hf (sqrt(2/3) sy + S0) / 2
            sy = sy0
            for _ in range(50):
                sy = yld(m, ep0 + np.sqrt(2 / 3) * g, np.sqrt(2 / 3) * g / dt,
                         T0 + 0.5 * g * hf * (np.sqrt(2 / 3) * sy + 0.0))
            return S_tr - np.sqrt(1.5) * G2 * g - sy

        g_ref = brentq(f, 1e-14, (S_tr - sy0) / (np.sqrt(1.5) * G2) * 1.0000001, xtol=1e-16, rtol=1e-14)
        assert out[9] == pytest.approx(g_ref, rel=1e-7, abs=1e-12)


def test_chung_hulbert_constants():
    """Explicit Chung-Hulbert on m a = -k x with the force taken at x_n
    (Explicit.C:438-536): (1 - aM) a1 + aM a0 = -w^2 x0,
    x1 = x0 + h v0 + h^2 ((1/2 - b) a0 + b a1), v1 = v0 + h ((1 - g) a0 + g a1).
    The optimal-dissipation constants make the amplification matrix's
    characteristic polynomial (lambda + rho_b)^3 at the bifurcation frequency
    (three conditions for aM, b, g); W_s is the stability limit at rho_b = 0
    and 1 and within 0.5 % of the true limit in between."""
    from scipy.optimize import brentq
    for rho_b in (0.0, 0.3, 0.5, 0.8182, 1.0):
        c = np.zeros(4)
        lib.oracle_chung_hulbert(ctypes.c_double(rho_b), P(c))
        aM, b, g, Ws = c

        def amp(W):
            a1 = np.array([-W ** 2, 0.0, -aM]) / (1 - aM)  # state (x, h v, h^2 a), w = 1, h = W
            x1 = np.array([1.0, 1.0, 0.5 - b]) + b * a1
            v1 = np.array([0.0, 1.0, 1 - g]) + g * a1
            return np.array([x1, v1, a1])

        Wb = brentq(lambda W: np.trace(amp(W)) + 3 * rho_b, 1e-6, Ws)
        np.testing.assert_allclose(np.poly(amp(Wb))[1:], [3 * rho_b, 3 * rho_b ** 2, rho_b ** 3], atol=1e-9)
        rad = lambda W: np.abs(np.linalg.eigvals(amp(W))).max()  # noqa: E731
        if rho_b in (0.0, 1.0):
            assert np.polyval(np.poly(amp(Ws)), -1.0) == pytest.approx(0.0, abs=1e-12)
        for W in np.linspace(0.05, 0.995 * Ws, 60):
            assert rad(W) <= 1.0 + 1e-9
    c = np.zeros(4)
    lib.oracle_chung_hulbert(ctypes.c_double(0.8182), P(c))
    np.testing.assert_allclose(c, [0.350016499835, 0.651521493113, 1.149983500165, 1.981818700333], atol=1e-12)


def char_len(elem_type, xyz):
    return lib.oracle_char_length(elem_type, P(np.ascontiguousarray(xyz, dtype=np.float64).ravel()))


def test_char_length_box_rectangle_parallelepiped():
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    assert char_len(4, base * [2.0, 3.0, 0.5]) == pytest.approx(0.5, rel=1e-14)
    # parallelepiped: L = volume / largest face area
    M = np.array([[2.0, 0.3, 0.1], [0.0, 1.5, 0.2], [0.0, 0.0, 1.0]])
    X = base @ M.T
    e1, e2, e3 = M[:, 0], M[:, 1], M[:, 2]
    vol = abs(np.dot(np.cross(e1, e2), e3))
    amax = max(np.linalg.norm(np.cross(e1, e2)), np.linalg.norm(np.cross(e2, e3)), np.linalg.norm(np.cross(e1, e3)))
    assert char_len(4, X) == pytest.approx(vol / amax, rel=1e-13)
    q = np.array([[0, 0, 0], [2.0, 0, 0], [2.0, 3.0, 0], [0, 3.0, 0]])
    assert char_len(3, q) == pytest.approx(6.0 / np.sqrt(13.0), rel=1e-14)


def test_ramp_functions():
    v = lambda t, x: lib.oracle_ramp_value(t, 0.0, 2.0, x)  # noqa: E731
    assert v(3, 1.0) == 0.5                 # Constant: 1 / (max - min)
    assert v(3, 2.5) == 0.0                 # outside the range
    assert v(1, 0.0) == pytest.approx(1.0)  # Linear: 2/mm (1 - x) -> integral 1
    for t in (1, 2, 3, 5):                  # each ramp integrates to 1 over [min, max]
        xs = np.linspace(0.0, 2.0, 20001)
        ys = np.array([v(t, x) for x in xs])
        assert np.trapezoid(ys, xs) == pytest.approx(1.0, rel=1e-3)
