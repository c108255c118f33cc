"""Pins the oracle's quaternion algebra, eigen-decomposition, body-index
cleaning and free-rotor rotations against independent mathematics."""
import numpy as np
import pytest

from quat import B, Bt, free_rotor_rk4, hamilton, random_unit_quaternion, rotation_matrix


def test_quat_from_matrix_reproduces_rotation(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = random_unit_quaternion(rng)
        M = rotation_matrix(q).T          # space -> body
        qq = oracle.quat_from_matrix(M)
        x = rng.standard_normal(3)
        assert np.allclose(oracle.quat_rotate(qq, x), M @ x, atol=1e-14)
        assert np.allclose(oracle.quat_rotate(qq, x, transpose=True), M.T @ x, atol=1e-14)
        assert abs(np.linalg.norm(qq) - 1) < 1e-15


def test_A_is_Bt_C_and_matches_rotation_matrix(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        q = random_unit_quaternion(rng)
        x = rng.standard_normal(3)
        assert np.allclose(oracle.quat_rotate(q, x, transpose=True), rotation_matrix(q) @ x, atol=1e-14)
        # B(q)^T B(q) = |q|^2 I and B(q)^T q = 0
        Bm = np.stack([B(q, e) for e in np.eye(3)], axis=1)
        assert np.allclose(Bm.T @ Bm, np.eye(3), atol=1e-14)
        assert np.allclose(Bt(q, q), 0, atol=1e-15)


def test_eigen_against_numpy(oracle):
    rng = np.random.default_rng(3)
    for _ in range(300):
        X = rng.standard_normal((3, 3))
        A = X @ X.T + np.diag(rng.uniform(0, 1, 3))
        w, V = oracle.eigen(A)
        ref = np.sort(np.linalg.eigvalsh(A))[::-1]
        assert np.allclose(w, ref, rtol=1e-12, atol=1e-13)
        assert w[0] >= w[1] >= w[2]
        assert np.allclose(V @ V.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(V) == pytest.approx(1.0, abs=1e-12)
        assert np.allclose(V.T @ np.diag(w) @ V, A, atol=1e-12 * np.abs(A).max())


def test_eigen_degenerate_and_diagonal(oracle):
    # diagonal input takes the p1 < eps branch; symmetric top takes the degenerate branch
    w, V = oracle.eigen(np.diag([1.0, 3.0, 2.0]))
    assert np.allclose(w, [3, 2, 1])
    R = rotation_matrix(random_unit_quaternion(np.random.default_rng(4)))
    A = R @ np.diag([2.0, 2.0, 1.0]) @ R.T
    w, V = oracle.eigen(A)
    assert np.allclose(w, [2, 2, 1], atol=1e-12)
    assert np.allclose(V.T @ np.diag(w) @ V, A, atol=1e-12)


def clean_spec(idx):
    """Brute-force statement of cleanBodyIndices (RigidBodySystem.cpp:28-49)."""
    labels = sorted(set(int(v) for v in idx if v > 0))
    rank = {v: i + 1 for i, v in enumerate(labels)}
    return np.array([rank[v] if v > 0 else 0 for v in idx], dtype=np.int32)


def test_clean_body_indices(oracle):
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 100, 1000):
        for _ in range(20):
            idx = rng.integers(-3, 40, size=n)
            assert np.array_equal(oracle.clean_body_indices(idx), clean_spec(idx))
    # unique positive labels become 1-atom bodies in the code (the comment says free)
    assert np.array_equal(oracle.clean_body_indices([0, 7, 3, 3]), [0, 2, 1, 1])


def random_top(rng, linear=False):
    I = np.sort(rng.uniform(0.004, 0.03, 3))[::-1]
    if linear:
        I = np.array([I[0], I[0], 0.0])
    invI = np.where(I > 0, 1.0 / np.where(I > 0, I, 1), 0.0)
    return I, invI


def test_uniaxial_closed_form(oracle):
    """A body spinning about principal axis 1 only: q(t) = q0 (x) (cos wt/2, sin wt/2 e1)."""
    rng = np.random.default_rng(6)
    for mode in (1, 10, 0):
        I, invI = random_top(rng)
        q0 = random_unit_quaternion(rng)
        L = np.array([0.2, 0.0, 0.0])
        pi = 2 * B(q0, L)
        dt = 0.01
        q, p = oracle.rotate(q0, pi, I, invI, 6, dt, mode)
        w = L[0] * invI[0]
        expect = hamilton(q0, np.array([np.cos(w * dt / 2), np.sin(w * dt / 2), 0, 0]))
        assert np.allclose(q, expect, atol=1e-14)
        assert np.allclose(p, 2 * B(expect, L), atol=1e-14)


@pytest.mark.parametrize("dt", [0.002, 0.05, 0.4, 3.0])
def test_exact_rotation_matches_rk4(oracle, dt):
    """Exact mode against RK4 integration of Euler's equations (both elliptic
    branches, half-period jumps at the larger dt)."""
    rng = np.random.default_rng(int(dt * 1000))
    for trial in range(5):
        I, invI = random_top(rng)
        q0 = random_unit_quaternion(rng)
        L = rng.standard_normal(3) * 0.15
        q, p = oracle.rotate(q0, 2 * B(q0, L), I, invI, 6, dt, 0)
        qr, Lr = free_rotor_rk4(q0, L, invI, dt, nsub=max(1500, int(dt * 3000)))
        assert np.allclose(q, qr, atol=5e-11), (trial, q, qr)
        assert np.allclose(Bt(q, p) / 2, Lr, atol=5e-11)


def test_nosquish_converges_to_rk4(oracle):
    rng = np.random.default_rng(8)
    I, invI = random_top(rng)
    q0 = random_unit_quaternion(rng)
    L = rng.standard_normal(3) * 0.15
    qr, Lr = free_rotor_rk4(q0, L, invI, 0.02)
    errs = []
    for n in (10, 20, 40):
        q, _ = oracle.rotate(q0, 2 * B(q0, L), I, invI, 6, 0.02, n)
        errs.append(np.abs(q - qr).max())
    assert errs[0] / errs[1] == pytest.approx(4.0, rel=0.2)   # second order in h
    assert errs[1] / errs[2] == pytest.approx(4.0, rel=0.2)


def test_rotation_invariants_and_reversibility(oracle):
    rng = np.random.default_rng(9)
    for mode in (0, 10):
        I, invI = random_top(rng)
        q0 = random_unit_quaternion(rng)
        L = rng.standard_normal(3) * 0.2
        pi0 = 2 * B(q0, L)
        q, p = oracle.rotate(q0, pi0, I, invI, 6, 0.01, mode)
        Lb = Bt(q, p) / 2
        # |q|, q.pi, body |L| and KE_rot conserved; space-frame L conserved
        assert abs(np.linalg.norm(q) - 1) < 1e-14
        assert abs(q @ p) < 1e-14
        assert np.linalg.norm(Lb) == pytest.approx(np.linalg.norm(L), rel=1e-13)
        # exact mode conserves KE_rot; NO-SQUISH (symplectic splitting) to O(h^2)
        assert Lb @ (invI * Lb) == pytest.approx(L @ (invI * L), rel=1e-12 if mode == 0 else 1e-3)
        assert np.allclose(rotation_matrix(q) @ Lb, rotation_matrix(q0) @ L, atol=1e-13)
        # time reversal: rotate by -dt returns to the start
        qb, pb = oracle.rotate(q, p, I, invI, 6, -0.01, mode)
        assert np.allclose(qb, q0, atol=1e-13) and np.allclose(pb, pi0, atol=1e-13)


def test_linear_rotor_policy(oracle):
    """dof-5 bodies in exact mode: the closed-form linear rotor (flag on)
    agrees with RK4; the reference's elliptic path divides by I0 - I1 = 0."""
    rng = np.random.default_rng(10)
    I, invI = random_top(rng, linear=True)
    q0 = random_unit_quaternion(rng)
    L = np.array([0.1, -0.05, 0.0])
    q, p = oracle.rotate(q0, 2 * B(q0, L), I, invI, 5, 0.05, 0, oracle.FIX_ALL)
    qr, Lr = free_rotor_rk4(q0, L, invI, 0.05)
    assert np.allclose(q, qr, atol=1e-11)
    qn, _ = oracle.rotate(q0, 2 * B(q0, L), I, invI, 5, 0.05, 0, 0)
    assert not np.all(np.isfinite(qn))
    # NO-SQUISH handles dof-5 bodies and converges to the same motion
    q10, _ = oracle.rotate(q0, 2 * B(q0, L), I, invI, 5, 0.05, 200, 0)
    assert np.allclose(q10, qr, atol=1e-7)
