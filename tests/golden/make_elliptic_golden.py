"""Writes tests/golden/elliptic.json: mpmath (40 digits) values of Carlson's
R_F, R_J, R_C, the Omega function of the exact rotation and Jacobi's
sn/cn/dn on a seeded grid of arguments, including the edge regions the exact
rotation reaches (x = 0 in R_F/R_J, m near 0 and 1, |u| beyond K).

    python tests/golden/make_elliptic_golden.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
rng = np.random.default_rng(20240601)
out = {"rf": [], "rj": [], "rc": [], "omega": [], "jacobi": []}

for _ in range(60):
    x, y, z = rng.uniform(0, 3, 3)
    out["rf"].append([x, y, z, float(mp.elliprf(x, y, z))])
for m in (0.0, 1e-12, 0.3, 0.9, 0.999999):
    out["rf"].append([0.0, 1.0 - m, 1.0, float(mp.elliprf(0, 1 - m, 1))])
for _ in range(60):
    x, y, z = rng.uniform(0, 3, 3)
    p = rng.uniform(0.05, 4)
    out["rj"].append([x, y, z, p, float(mp.elliprj(x, y, z, p))])
for x, y, z, p in ((0.0, 0.5, 1.0, 1.5), (0.0, 1e-3, 1.0, 3.0), (1.0, 1.0, 1.0, 1.0)):
    out["rj"].append([x, y, z, p, float(mp.elliprj(x, y, z, p))])
for _ in range(40):
    x = rng.uniform(0, 3)
    y = rng.uniform(0.01, 3)
    out["rc"].append([x, y, float(mp.elliprc(x, y))])
out["rc"].append([1.0, 1.0, 1.0])
for _ in range(60):
    x = rng.uniform(-1, 1)
    n = rng.uniform(0, 5)
    m = rng.uniform(0, 1)
    x2 = mp.mpf(x) ** 2
    val = -(mp.mpf(n) * mp.mpf(x) ** 3 / 3) * mp.elliprj(1 - x2, 1 - m * x2, 1, 1 + n * x2)
    out["omega"].append([x, n, m, float(val)])
for x, n, m in ((1.0, 0.7, 0.4), (-1.0, 2.0, 0.9), (0.0, 1.0, 0.5)):
    x2 = mp.mpf(x) ** 2
    val = -(mp.mpf(n) * mp.mpf(x) ** 3 / 3) * mp.elliprj(1 - x2, 1 - m * x2, 1, 1 + n * x2)
    out["omega"].append([x, n, m, float(val)])
for _ in range(80):
    u = rng.uniform(-12, 12)
    m = rng.uniform(0, 1)
    out["jacobi"].append([u, m] + [float(mp.ellipfun(k, u, m=m)) for k in ("sn", "cn", "dn")])
for u, m in ((0.7, 0.0), (0.7, 1e-17), (1.3, 1.0), (-2.0, 1.0 - 1e-17), (3.0, 0.999999), (float(mp.ellipk(0.5)), 0.5)):
    out["jacobi"].append([u, m] + [float(mp.ellipfun(k, u, m=m)) for k in ("sn", "cn", "dn")])

path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elliptic.json")
with open(path, "w") as f:
    json.dump(out, f, indent=0)
print("wrote", path, {k: len(v) for k, v in out.items()})
