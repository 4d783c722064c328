"""Pins the CPU oracle's element routines (CPU only).

Each check is chosen so that a plausible mistake in the oracle (dropped term,
wrong sign or index, transposed operand) fails it:
  * golden vectors from the reference's own element code (tests/golden);
  * an independent numpy derivation (3x3x3 Gauss / barycentric gradients) on
    affine elements, where 2x2x2 Gauss is exact;
  * invariants: symmetry, rigid-body null space, PSD rank, scaling, sign flip;
  * TL: u = 0 reduces to the linear Ke, the tangent is the central finite
    difference of the internal force, a rigid rotation gives zero force.
"""
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

from conftest import rel_max
from oracle import oracle as O

E, NU = 210000.0, 0.3
HEX0 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
TET0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
SGN = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def stiffness_tensor(E, nu):
    lam, mu = lame(E, nu)
    d = np.eye(3)
    return lam * np.einsum("ij,kl->ijkl", d, d) + mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d))


def hex_ke_independent(X, E, nu, order=3):
    """Ke_acbd = int dN_a/dx_j C_cjdl dN_b/dx_l with tensor C and order^3 Gauss."""
    Ct = stiffness_tensor(E, nu)
    x, w = leggauss(order)
    K = np.zeros((8, 3, 8, 3))
    for i, r in enumerate(x):
        for j, s in enumerate(x):
            for k, t in enumerate(x):
                dN = np.stack([SGN[:, 0] * (1 + SGN[:, 1] * s) * (1 + SGN[:, 2] * t),
                               SGN[:, 1] * (1 + SGN[:, 0] * r) * (1 + SGN[:, 2] * t),
                               SGN[:, 2] * (1 + SGN[:, 0] * r) * (1 + SGN[:, 1] * s)]) / 8.0
                J = dN @ X
                g = np.linalg.solve(J, dN)  # 3 x 8 physical gradients
                K += w[i] * w[j] * w[k] * np.linalg.det(J) * np.einsum("ja,cjdl,lb->acbd", g, Ct, g)
    return K.reshape(24, 24)


def tet_ke_independent(X, E, nu):
    """Barycentric gradients from inv([1 x y z]); Ke = V * g C g (signed V)."""
    A = np.hstack([np.ones((4, 1)), X])
    G = np.linalg.inv(A)[1:, :]  # 3 x 4: grad of barycentric coordinate a
    V = np.linalg.det(A) / 6.0
    return V * np.einsum("ja,cjdl,lb->acbd", G, stiffness_tensor(E, nu), G).reshape(12, 12)


def rigid_modes(X):
    n = len(X)
    R = np.zeros((3 * n, 6))
    for a in range(n):
        x, y, z = X[a]
        R[3 * a:3 * a + 3, :3] = np.eye(3)
        R[3 * a:3 * a + 3, 3] = [-y, x, 0]
        R[3 * a:3 * a + 3, 4] = [0, -z, y]
        R[3 * a:3 * a + 3, 5] = [z, 0, -x]
    return R


def test_dmat_closed_form():
    D = O.dmat(E, NU)
    lam, mu = lame(E, NU)
    assert np.allclose(D[:3, :3], lam + 2 * mu * np.eye(3), rtol=1e-15)
    assert np.allclose(np.diag(D)[3:], mu, rtol=1e-15)
    assert np.count_nonzero(D) == 12


def test_hex_ke_matches_reference_vectors(ref_vectors):
    for X, K in zip(ref_vectors["hex_X"], ref_vectors["hex_Ke"]):
        assert rel_max(O.c3d8_ke(X, E, NU), K) < 1e-13


def test_tet_ke_matches_reference_vectors(ref_vectors):
    for X, K in zip(ref_vectors["tet_X"], ref_vectors["tet_Ke"]):
        assert rel_max(O.c3d4_ke(X, E, NU), K) < 1e-13


@pytest.mark.parametrize("A", [np.eye(3), np.diag([2.0, 0.5, 1.5]), np.array([[1, 0.3, 0], [0, 1, 0.2], [0.1, 0, 1.0]])])
def test_hex_ke_affine_independent(A):
    X = HEX0 @ A.T + np.array([3.0, -1.0, 2.0])
    assert rel_max(O.c3d8_ke(X, E, NU), hex_ke_independent(X, E, NU)) < 1e-13


def test_tet_ke_independent():
    rng = np.random.default_rng(3)
    for _ in range(5):
        X = TET0 + rng.uniform(-0.2, 0.2, TET0.shape)
        assert rel_max(O.c3d4_ke(X, E, NU), tet_ke_independent(X, E, NU)) < 1e-13


def test_ke_invariants():
    rng = np.random.default_rng(5)
    X = HEX0 + rng.uniform(-0.2, 0.2, HEX0.shape)
    K = O.c3d8_ke(X, E, NU)
    assert rel_max(K, K.T) < 1e-14
    assert np.abs(K @ rigid_modes(X)).max() < 1e-10 * np.abs(K).max()
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert np.sum(ev > 1e-8 * ev.max()) == 18 and ev.min() > -1e-9 * ev.max()
    assert rel_max(O.c3d8_ke(2.5 * X, E, NU), 2.5 * K) < 1e-13          # Ke(aX) = a Ke(X)
    assert rel_max(O.c3d8_ke(X, 3 * E, NU), 3 * K) < 1e-14              # Ke linear in E
    # mirrored node numbering: detJ < 0 at every point, Ke changes sign
    perm = [4, 5, 6, 7, 0, 1, 2, 3]
    Km = O.c3d8_ke(X[perm], E, NU)
    P = np.zeros((24, 24))
    for new, old in enumerate(perm):
        P[3 * new:3 * new + 3, 3 * old:3 * old + 3] = np.eye(3)
    assert rel_max(Km, -(P @ K @ P.T)) < 1e-13
    T = O.c3d4_ke(TET0 + 0.1, E, NU)
    Ti = O.c3d4_ke((TET0 + 0.1)[[0, 2, 1, 3]], E, NU)
    Q = np.zeros((12, 12))
    for new, old in enumerate([0, 2, 1, 3]):
        Q[3 * new:3 * new + 3, 3 * old:3 * old + 3] = np.eye(3)
    assert rel_max(Ti, -(Q @ T @ Q.T)) < 1e-14


def test_mass_matches_reference_and_totals(ref_vectors):
    rho = float(ref_vectors["rho"])
    for X, m in zip(ref_vectors["hex_X"], ref_vectors["hex_mass"]):
        assert rel_max(O.c3d8_mass(X, rho), m) < 1e-13
    for X, m in zip(ref_vectors["tet_X"], ref_vectors["tet_mass"]):
        assert rel_max(O.c3d4_mass(X, rho), m) < 1e-13
    assert np.allclose(O.c3d8_mass(HEX0, 8.0), 1.0, rtol=1e-14)      # rho V / 8 per node
    assert np.allclose(O.c3d4_mass(TET0, 24.0), 1.0, rtol=1e-14)     # |V| rho / 4


def test_in_force_is_minus_ke_du(ref_vectors):
    """cal_in_force subtracts Ke du from the running force (src/Elements/C3D8.cpp:158-193)."""
    for X, du, f in zip(ref_vectors["hex_X"], ref_vectors["hex_u"], ref_vectors["hex_in_force"]):
        assert rel_max(-O.c3d8_ke(X, E, NU) @ du, f) < 1e-13
    for X, du, f in zip(ref_vectors["tet_X"], ref_vectors["tet_u"], ref_vectors["tet_in_force"]):
        assert rel_max(-O.c3d4_ke(X, E, NU) @ du, f) < 1e-13


def test_tl_matches_reference_vectors(ref_vectors):
    for X, u, K, f in zip(ref_vectors["hex_X"], ref_vectors["hex_u"], ref_vectors["hex_KT"], ref_vectors["hex_fint"]):
        Ko, fo = O.c3d8_nl(X, u, E, NU)
        assert rel_max(Ko, K) < 1e-13 and rel_max(fo, f) < 1e-12
    for X, u, K, f in zip(ref_vectors["tet_X"], ref_vectors["tet_u"], ref_vectors["tet_KT"], ref_vectors["tet_fint"]):
        Ko, fo = O.c3d4_nl(X, u, E, NU)
        assert rel_max(Ko, K) < 1e-13 and rel_max(fo, f) < 1e-12


@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_tl_consistency(kind):
    rng = np.random.default_rng(11)
    X = (HEX0 if kind == "hex" else TET0) + rng.uniform(-0.15, 0.15, (8 if kind == "hex" else 4, 3))
    nl = O.c3d8_nl if kind == "hex" else O.c3d4_nl
    ke = O.c3d8_ke if kind == "hex" else O.c3d4_ke
    nc = X.size
    K0, f0 = nl(X, np.zeros(nc), E, NU)
    assert rel_max(K0, ke(X, E, NU)) < 1e-14 and np.abs(f0).max() == 0.0
    u = rng.normal(0, 0.03, nc)
    K, f = nl(X, u, E, NU)
    eps = 1e-6
    Kfd = np.zeros((nc, nc))
    for j in range(nc):
        du = np.zeros(nc)
        du[j] = eps
        Kfd[:, j] = (nl(X, u + du, E, NU)[1] - nl(X, u - du, E, NU)[1]) / (2 * eps)
    assert rel_max(K, Kfd) < 1e-7
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    urot = (X @ R.T - X).ravel()
    assert np.abs(nl(X, urot, E, NU)[1]).max() < 1e-9 * E
