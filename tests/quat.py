"""Small numpy quaternion helpers for tests (reference convention, MatVec.cpp)."""
import numpy as np


def B(q, x):
    return np.array([-q[1] * x[0] - q[2] * x[1] - q[3] * x[2], q[0] * x[0] - q[3] * x[1] + q[2] * x[2],
                     q[3] * x[0] + q[0] * x[1] - q[1] * x[2], -q[2] * x[0] + q[1] * x[1] + q[0] * x[2]])


def Bt(q, y):
    return np.array([-q[1] * y[0] + q[0] * y[1] + q[3] * y[2] - q[2] * y[3],
                     -q[2] * y[0] - q[3] * y[1] + q[0] * y[2] + q[1] * y[3],
                     -q[3] * y[0] + q[2] * y[1] - q[1] * y[2] + q[0] * y[3]])


def hamilton(a, b):
    """Hamilton product a (x) b, scalar first."""
    a0, av = a[0], np.asarray(a[1:])
    b0, bv = b[0], np.asarray(b[1:])
    return np.concatenate([[a0 * b0 - av @ bv], a0 * bv + b0 * av + np.cross(av, bv)])


def rotation_matrix(q):
    """Body -> space matrix A(q)^T for q = (w, x, y, z)."""
    w, x, y, z = q
    return np.array([[w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]])


def free_rotor_rk4(q, L, invI, t, nsub=4000):
    """Free rigid body by RK4: dq/dt = q (x) (0, w) / 2, dL/dt = L x w, w = invI * L (body frame)."""
    q = np.array(q, dtype=float)
    L = np.array(L, dtype=float)
    h = t / nsub

    def f(q, L):
        w = invI * L
        return 0.5 * hamilton(q, np.concatenate([[0.0], w])), np.cross(L, w)

    for _ in range(nsub):
        k1 = f(q, L)
        k2 = f(q + 0.5 * h * k1[0], L + 0.5 * h * k1[1])
        k3 = f(q + 0.5 * h * k2[0], L + 0.5 * h * k2[1])
        k4 = f(q + h * k3[0], L + h * k3[1])
        q = q + h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        L = L + h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
    return q, L


def random_unit_quaternion(rng):
    q = rng.standard_normal(4)
    return q / np.linalg.norm(q)
