/*
 * Exports the reference's own elliptic routines for cross-checking the
 * oracle's independently written ones.  This file contains no reference code:
 * it includes the reference header in place (oracle/Makefile puts the
 * reference's openmmapi/include on the include path and writes only into
 * oracle/_ref/).  The header needs only <math.h>, <limits> and <algorithm>,
 * so it compiles from its own source with no stand-ins.
 * Test infrastructure only.
 */
#include "internal/ellipticFunctions.h"

extern "C" {
double ref_carlson_rf(double x, double y, double z) { return carlsonRF(x, y, z); }
double ref_carlson_rj(double x, double y, double z, double p) { return carlsonRJ(x, y, z, p); }
double ref_carlson_rc(double x, double y) { return carlsonRC(x, y); }
double ref_omega(double x, double n, double m) { return Omega(x, n, m); }
void ref_jacobi(double u, double m, double* sn, double* cn, double* dn) { jacobi(u, m, *sn, *cn, *dn); }
}
