"""Prints the Taylor-coefficient tables of the Jacobi elliptic functions used
by rbhip_device.h (kSnCoef, kCnCoef, kDnCoef).

sn, cn, dn(u | m) satisfy sn' = cn dn, cn' = -sn dn, dn' = -m sn cn with
sn(0) = 0, cn(0) = dn(0) = 1 (DLMF 22.13.1-3); the Cauchy products of these
recurrences give each Taylor coefficient exactly, as a polynomial in m with
rational coefficients.  The tables hold the odd powers 1..17 of sn and the
even powers 0..16 of cn and dn, each row a polynomial in m (ascending).

Usage: python gen_jacobi_series.py > tables.inc
"""
from fractions import Fraction


def _add(a, b):
    n = max(len(a), len(b))
    return [(a[i] if i < len(a) else 0) + (b[i] if i < len(b) else 0) for i in range(n)]


def _mul(a, b):
    r = [Fraction(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] += x * y
    return r


def coefficients(order):
    s = [[Fraction(0)] for _ in range(order + 1)]
    c = [[Fraction(0)] for _ in range(order + 1)]
    d = [[Fraction(0)] for _ in range(order + 1)]
    c[0] = [Fraction(1)]
    d[0] = [Fraction(1)]
    for k in range(order):
        cd, sd, sc = [Fraction(0)], [Fraction(0)], [Fraction(0)]
        for i in range(k + 1):
            cd = _add(cd, _mul(c[i], d[k - i]))
            sd = _add(sd, _mul(s[i], d[k - i]))
            sc = _add(sc, _mul(s[i], c[k - i]))
        s[k + 1] = [x / (k + 1) for x in cd]
        c[k + 1] = [-x / (k + 1) for x in sd]
        d[k + 1] = [-x / (k + 1) for x in [Fraction(0)] + sc]   # the factor m
    return s, c, d


def table(name, polys, powers, width):
    rows = []
    for k in powers:
        p = list(polys[k]) + [Fraction(0)] * (width - len(polys[k]))
        rows.append("    {" + ", ".join(repr(float(x)) for x in p) + "},")
    return f"__device__ constexpr double {name}[{len(powers)}][{width}] = {{\n" + "\n".join(rows) + "\n};"


def main(order=17):
    s, c, d = coefficients(order)
    width = order // 2 + 1
    print(table("kSnCoef", s, range(1, order + 1, 2), width))
    print(table("kCnCoef", c, range(0, order, 2), width))
    print(table("kDnCoef", d, range(0, order, 2), width))


if __name__ == "__main__":
    main()
