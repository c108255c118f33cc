"""The Taylor-series tables of the Jacobi elliptic functions in the device
code (kSnCoef, kCnCoef, kDnCoef in rbhip_device.h, used by the exact rotation
for short steps): the header holds what gen_jacobi_series.py generates, and
the truncated series agree with mpmath's sn, cn, dn to 2e-16 for |u| <= 0.2
and 0 <= m < 1 (including m within 1e-12 of 0 and of 1)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "openmm_rigidbody_plugin_amd", "csrc")


def header_tables():
    text = open(os.path.join(CSRC, "rbhip_device.h")).read()
    out = {}
    for name in ("kSnCoef", "kCnCoef", "kDnCoef"):
        body = re.search(name + r"\[9\]\[9\] = \{(.*?)\n\};", text, re.S).group(1)
        rows = re.findall(r"\{([^{}]*)\}", body)
        out[name] = np.array([[float(x) for x in r.split(",")] for r in rows])
    return out


def test_header_tables_are_generated():
    gen = subprocess.run([sys.executable, os.path.join(CSRC, "gen_jacobi_series.py")], capture_output=True,
                         text=True, check=True).stdout
    text = open(os.path.join(CSRC, "rbhip_device.h")).read()
    for line in gen.strip().splitlines():
        assert line in text, line


def series(u, m, t):
    pm = lambda row: np.polynomial.polynomial.polyval(m, row)
    u2 = u * u
    sn = u * sum(pm(t["kSnCoef"][k]) * u2 ** k for k in range(9))
    cn = sum(pm(t["kCnCoef"][k]) * u2 ** k for k in range(9))
    dn = sum(pm(t["kDnCoef"][k]) * u2 ** k for k in range(9))
    return sn, cn, dn


def test_series_match_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    t = header_tables()
    rng = np.random.default_rng(5)
    worst = 0.0
    for i in range(400):
        m = [rng.uniform(0, 1), 1 - 10 ** rng.uniform(-12, -1), 10 ** rng.uniform(-12, -1)][i % 3]
        u = rng.uniform(-0.2, 0.2)
        ref = [float(mp.ellipfun(f, u, m=m)) for f in ("sn", "cn", "dn")]
        got = series(u, m, t)
        worst = max(worst, max(abs(g - r) for g, r in zip(got, ref)))
    assert worst <= 4e-16, worst


def test_series_identities():
    t = header_tables()
    rng = np.random.default_rng(6)
    for _ in range(200):
        m, u = rng.uniform(0, 1), rng.uniform(-0.2, 0.2)
        sn, cn, dn = series(u, m, t)
        assert abs(sn * sn + cn * cn - 1) <= 7e-16   # 3 ulp of 1: rounding of the sums
        assert abs(dn * dn + m * sn * sn - 1) <= 7e-16
