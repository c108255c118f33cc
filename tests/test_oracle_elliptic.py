"""Pins the oracle's elliptic routines (Carlson 1995, DLMF 22.20) three ways:
mpmath golden vectors, the reference's own ellipticFunctions.h compiled into
oracle/_ref (when present), and closed forms / identities."""
import json
import math
import os

import numpy as np
import pytest
import scipy.special as sps

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "elliptic.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_carlson_rf_golden(oracle, golden):
    for x, y, z, val in golden["rf"]:
        assert oracle.carlson_rf(x, y, z) == pytest.approx(val, rel=2e-15)


def test_carlson_rj_golden(oracle, golden):
    for x, y, z, p, val in golden["rj"]:
        assert oracle.carlson_rj(x, y, z, p) == pytest.approx(val, rel=4e-15)


def test_carlson_rc_golden(oracle, golden):
    for x, y, val in golden["rc"]:
        assert oracle.carlson_rc(x, y) == pytest.approx(val, rel=2e-15)


def test_omega_golden(oracle, golden):
    for x, n, m, val in golden["omega"]:
        assert oracle.omega(x, n, m) == pytest.approx(val, rel=4e-15, abs=1e-300)


def test_jacobi_golden(oracle, golden):
    for u, m, sn, cn, dn in golden["jacobi"]:
        s, c, d = oracle.jacobi(u, m)
        assert abs(s - sn) < 6e-15 and abs(c - cn) < 6e-15 and abs(d - dn) < 6e-15, (u, m)


def test_closed_forms(oracle):
    # R_F(0, 1-m, 1) = K(m); R_C(x, y) = arccos(sqrt(x/y)) / sqrt(y - x) for x < y
    for m in (0.0, 0.2, 0.5, 0.9, 0.999):
        assert oracle.carlson_rf(0.0, 1.0 - m, 1.0) == pytest.approx(sps.ellipk(m), rel=1e-14)
    for x, y in ((0.2, 1.0), (1.0, 4.0), (0.0, 2.0)):
        assert oracle.carlson_rc(x, y) == pytest.approx(math.acos(math.sqrt(x / y)) / math.sqrt(y - x), rel=1e-14)
    # R_J(x, y, z, z) = R_D(x, y, z) and R_J(x, x, x, x) = x^(-3/2)
    assert oracle.carlson_rj(1.0, 1.0, 1.0, 1.0) == pytest.approx(1.0, rel=1e-15)
    assert oracle.carlson_rj(4.0, 4.0, 4.0, 4.0) == pytest.approx(0.125, rel=1e-15)
    # Jacobi identities and scipy
    rng = np.random.default_rng(7)
    for _ in range(50):
        u, m = rng.uniform(-10, 10), rng.uniform(0, 1)
        s, c, d = oracle.jacobi(u, m)
        assert s * s + c * c == pytest.approx(1.0, abs=4e-15)
        assert d * d + m * s * s == pytest.approx(1.0, abs=4e-15)
        ss, cc, dd, _ = sps.ellipj(u, m)
        assert abs(s - ss) < 1e-13 and abs(c - cc) < 1e-13 and abs(d - dd) < 1e-13


def test_against_reference_header(oracle):
    """The reference's own routines (oracle/_ref, built from its header)."""
    ref = oracle.reference_elliptic()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes
    rng = np.random.default_rng(11)
    for _ in range(300):
        x, y, z = rng.uniform(0, 2, 3)
        p = rng.uniform(0.1, 3)
        assert oracle.carlson_rf(x, y, z) == pytest.approx(ref.ref_carlson_rf(x, y, z), rel=2e-15)
        assert oracle.carlson_rj(x, y, z, p) == pytest.approx(ref.ref_carlson_rj(x, y, z, p), rel=4e-15)
        assert oracle.carlson_rc(x, y + 0.01) == pytest.approx(ref.ref_carlson_rc(x, y + 0.01), rel=2e-15)
        xx, nn, mm = rng.uniform(-1, 1), rng.uniform(0, 3), rng.uniform(0, 1)
        assert oracle.omega(xx, nn, mm) == pytest.approx(ref.ref_omega(xx, nn, mm), rel=4e-15, abs=1e-300)
        u = rng.uniform(-8, 8)
        sn, cn, dn = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ref.ref_jacobi(u, mm, sn, cn, dn)
        s, c, d = oracle.jacobi(u, mm)
        assert abs(s - sn.value) < 1e-14 and abs(c - cn.value) < 1e-14 and abs(d - dn.value) < 1e-14


def test_planted_mistake_is_caught(golden):
    """A plausible restatement slip (dropping the E3 term of R_F's series)
    misses the golden vectors: the check has teeth."""
    def rf_wrong(x, y, z):
        A0 = (x + y + z) / 3
        Q = (3e-16) ** (-1 / 6) * max(abs(A0 - x), abs(A0 - y), abs(A0 - z))
        xm, ym, zm, Am, sc = x, y, z, A0, 1.0
        while sc * Q >= abs(Am):
            lam = math.sqrt(xm) * math.sqrt(ym) + math.sqrt(xm) * math.sqrt(zm) + math.sqrt(ym) * math.sqrt(zm)
            xm, ym, zm, Am, sc = (xm + lam) / 4, (ym + lam) / 4, (zm + lam) / 4, (Am + lam) / 4, sc / 4
        X, Y = (A0 - x) * sc / Am, (A0 - y) * sc / Am
        Z = -(X + Y)
        E2 = X * Y - Z * Z
        return (1 - E2 / 10 + E2 * E2 / 24) / math.sqrt(Am)
    bad = sum(abs(rf_wrong(x, y, z) / v - 1) > 2e-15 for x, y, z, v in golden["rf"])
    assert bad > 0
