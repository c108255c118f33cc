/*
 * rbref.cpp — CPU FP64 restatement of the reference RigidBodyIntegrator path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Built with g++ -ffp-contract=off so
 * every operation is an IEEE double op in the order the reference writes it.
 *
 * Function → reference source it follows → how it is pinned (tests/):
 *   Vec3/Mat3/Quat ops      MatVec.cpp:15-564, OpenMM Vec3 (external)
 *                           → test_oracle_math: A(Quat(M))x == Mx, B/C
 *                             orthogonality, Bt(q)B(q) = I
 *   clean_body_indices      RigidBodySystem.cpp:28-49
 *                           → brute-force spec on random index vectors
 *   RigidBodySystem::init   RigidBodySystem.cpp:55-114
 *                           → brute-force atom ordering spec
 *   eigenvalues (Smith)     eigenDecomposition.cpp:72-100
 *   eigenvectors (Kopp)     eigenDecomposition.cpp:104-176
 *                           → numpy.linalg.eigh, reconstruction V^T W V = A
 *   quat_from_matrix        MatVec.cpp:344-373 (Shepperd 1978)
 *                           → A(q)x == Mx for random rotations
 *   carlson_rf/rj/rc        written from Carlson, Numer. Algorithms 10 (1995)
 *                           13-26 (Algorithms 1-3), not from the reference
 *                           header → mpmath golden vectors + the reference
 *                           header itself compiled in oracle/_ref
 *   jacobi                  written from DLMF 22.20(ii) (AGM/descending
 *                           Landen) → mpmath.ellipfun golden vectors + _ref
 *   omega                   ellipticFunctions.h:251-254 (definition)
 *                           → mpmath quadrature golden vectors
 *   buildGeometry           RigidBody.cpp:65-116 → rigidity, d = A delta,
 *                           inertia reconstruction
 *   buildDynamics           RigidBody.cpp:123-142 → KE identity
 *                           (sum 1/2 m v^2 == KEt + KEr for rigid velocities)
 *   uniaxial / NO-SQUISH    RigidBody.cpp:189-231 → closed-form single-axis
 *                           rotation, reversibility, |q|, q.pi = 0
 *   exactRotation           RigidBody.cpp:241-308 → free asymmetric top vs
 *                           RK4 integration of Euler's equations (mpmath-free,
 *                           tight step), L and KE conservation, NO-SQUISH
 *                           n->large convergence
 *   updateAtomicPositions   RigidBody.cpp:148-153 → rigidity
 *   updateAtomicVelocities  RigidBody.cpp:159-168 → test_oracle_velocities:
 *                             V_j = central difference of r_j(t), V_j - v_cm =
 *                             w e_z x (r_j - c) for a spinning dimer, constrained
 *                             triangle velocities -> rigid-body velocities as
 *                             dt -> 0; each fails with the planted delta x w
 *   forceAndTorque          RigidBody.cpp:174-183 → brute-force sums
 *   integratePart1/2        RigidBodySystem.cpp:170-204
 *   execute                 ReferenceRigidBodyKernels.cpp:82-108
 *                           → testSingleBond KAT (tests/TestRigidBodyIntegrator.h:49-83),
 *                             energy conservation, time reversibility
 *   computeKineticEnergies  RigidBodySystem.cpp:210-220 → KE identity
 *   refined energies        the CUDA platform's COMPMOD terms,
 *                           rigidbodyintegrator.cu:270-292, 318-320, 380-383,
 *                           433-469, CudaRigidBodyKernels.cpp:405-494 (the
 *                           reference platform returns plain values)
 *                           → test_oracle_refined: exact at constant velocity,
 *                             third order for free rotors, shadow-energy
 *                             fluctuation O(dt^3) vs O(dt^2), closed-form U_ref
 *   free-atom constraints   the call sites ReferenceRigidBodyKernels.cpp:92-104,
 *                           the body-atom refusal RigidBodySystem.cpp:107-112
 *                           and the DOF count :130; the solver (OpenMM's, not
 *                           in the reference tree) as SHAKE + RATTLE
 *                           → test_oracle_constraints: the projection equals a
 *                             Newton solve of the constraint equations, RATTLE
 *                             velocities orthogonal to the bonds, constrained
 *                             free atoms converge to the rigid-body trajectory
 *                             as dt -> 0, energy conservation, the reference's
 *                             (commented-out) constraint tests
 *
 * Parity unpinned (no check fixes them beyond the restatement itself): the
 * bit-level FP64 output of the original exactRotation composition and of the
 * degenerate eigenvector branches (eigenDecomposition.cpp:138-171), because
 * the reference cannot be compiled here (it needs OpenMM headers).
 */
#include "rbref.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

const double kEps = std::numeric_limits<double>::epsilon();
const double kPi = 3.14159265358979323846264338328;

/* ------------------------------------------------------------------------ */
/* Vec3: the componentwise algebra of OpenMM::Vec3 (division multiplies by the
   reciprocal, as OpenMM's Vec3::operator/ does).                            */
/* ------------------------------------------------------------------------ */
struct Vec3 {
    double c[3];
    Vec3() : c{0.0, 0.0, 0.0} {}
    Vec3(double x, double y, double z) : c{x, y, z} {}
    double operator[](int i) const { return c[i]; }
    double& operator[](int i) { return c[i]; }
    Vec3 operator+(const Vec3& o) const { return Vec3(c[0] + o[0], c[1] + o[1], c[2] + o[2]); }
    Vec3 operator-(const Vec3& o) const { return Vec3(c[0] - o[0], c[1] - o[1], c[2] - o[2]); }
    Vec3 operator-() const { return Vec3(-c[0], -c[1], -c[2]); }
    Vec3 operator*(double s) const { return Vec3(c[0] * s, c[1] * s, c[2] * s); }
    Vec3 operator/(double s) const { double r = 1.0 / s; return Vec3(c[0] * r, c[1] * r, c[2] * r); }
    Vec3& operator+=(const Vec3& o) { c[0] += o[0]; c[1] += o[1]; c[2] += o[2]; return *this; }
    Vec3& operator-=(const Vec3& o) { c[0] -= o[0]; c[1] -= o[1]; c[2] -= o[2]; return *this; }
    Vec3& operator*=(double s) { c[0] *= s; c[1] *= s; c[2] *= s; return *this; }
    Vec3& operator/=(double s) { double r = 1.0 / s; c[0] *= r; c[1] *= r; c[2] *= r; return *this; }
    double dot(const Vec3& o) const { return c[0] * o[0] + c[1] * o[1] + c[2] * o[2]; }
    Vec3 cross(const Vec3& o) const {
        return Vec3(c[1] * o[2] - c[2] * o[1], c[2] * o[0] - c[0] * o[2], c[0] * o[1] - c[1] * o[0]);
    }
};

/* Diag3 * Vec3 (MatVec.cpp:297-300) */
inline Vec3 diagMul(const Vec3& d, const Vec3& x) { return Vec3(d[0] * x[0], d[1] * x[1], d[2] * x[2]); }

/* Mat3 stored by rows (MatVec.cpp:15-188) */
struct Mat3 {
    Vec3 r[3];
    Mat3() {}
    Mat3(const Vec3& a, const Vec3& b, const Vec3& c) { r[0] = a; r[1] = b; r[2] = c; }
    const Vec3& operator[](int i) const { return r[i]; }
    Vec3& operator[](int i) { return r[i]; }
    Vec3 col(int j) const { return Vec3(r[0][j], r[1][j], r[2][j]); }
    Vec3 diag() const { return Vec3(r[0][0], r[1][1], r[2][2]); }
    Mat3 t() const { return Mat3(col(0), col(1), col(2)); }
    Mat3& operator+=(const Mat3& o) { r[0] += o[0]; r[1] += o[1]; r[2] += o[2]; return *this; }
    Mat3 operator*(double s) const { return Mat3(r[0] * s, r[1] * s, r[2] * s); }
    Vec3 operator*(const Vec3& x) const { return Vec3(r[0].dot(x), r[1].dot(x), r[2].dot(x)); }
    Mat3 minusDiag(double q) const {           /* Mat3 - Diag3(q), MatVec.cpp:93-100 */
        Mat3 a = *this; a[0][0] -= q; a[1][1] -= q; a[2][2] -= q; return a;
    }
    Mat3 symmetric() const {                   /* MatVec.cpp:141-147 */
        Mat3 a = *this; a[1][0] = a[0][1]; a[2][0] = a[0][2]; a[2][1] = a[1][2]; return a;
    }
    double det() const {                       /* MatVec.cpp:156-164 */
        const Vec3 &a0 = r[0], &a1 = r[1], &a2 = r[2];
        return a0[0] * (a1[1] * a2[2] - a2[1] * a1[2]) -
               a0[1] * (a1[0] * a2[2] - a2[0] * a1[2]) +
               a0[2] * (a1[0] * a2[1] - a2[0] * a1[1]);
    }
};

/* Projection(x) = |x|^2 I - x x^T, in the order of MatVec.cpp:555-564 */
Mat3 projection(const Vec3& x) {
    Mat3 p((-x) * x[0], (-x) * x[1], (-x) * x[2]);
    double xtx = x.dot(x);
    p[0][0] += xtx; p[1][1] += xtx; p[2][2] += xtx;
    return p;
}

/* Quaternion with the B/C matrix products of MatVec.cpp:495-549 */
struct Quat {
    double c[4];
    Quat() : c{0.0, 0.0, 0.0, 0.0} {}
    Quat(double a, double b, double d, double e) : c{a, b, d, e} {}
    double operator[](int i) const { return c[i]; }
    double& operator[](int i) { return c[i]; }
    Quat operator+(const Quat& o) const { return Quat(c[0] + o[0], c[1] + o[1], c[2] + o[2], c[3] + o[3]); }
    Quat operator-(const Quat& o) const { return Quat(c[0] - o[0], c[1] - o[1], c[2] - o[2], c[3] - o[3]); }
    Quat operator*(double s) const { return Quat(c[0] * s, c[1] * s, c[2] * s, c[3] * s); }
    Quat& operator+=(const Quat& o) { for (int i = 0; i < 4; i++) c[i] += o[i]; return *this; }
    Quat& operator/=(double s) { double r = 1.0 / s; for (int i = 0; i < 4; i++) c[i] *= r; return *this; }
    double dot(const Quat& o) const { return c[0] * o[0] + c[1] * o[1] + c[2] * o[2] + c[3] * o[3]; }
    double norm() const { return std::sqrt(dot(*this)); }
    /* B(q) x: columns are q (x) e_k for k = 1, 2, 3 */
    Quat B(const Vec3& x) const {
        return Quat(-c[1] * x[0] - c[2] * x[1] - c[3] * x[2],
                     c[0] * x[0] - c[3] * x[1] + c[2] * x[2],
                     c[3] * x[0] + c[0] * x[1] - c[1] * x[2],
                    -c[2] * x[0] + c[1] * x[1] + c[0] * x[2]);
    }
    /* C(q) x: columns are e_k (x) q */
    Quat C(const Vec3& x) const {
        return Quat(-c[1] * x[0] - c[2] * x[1] - c[3] * x[2],
                     c[0] * x[0] + c[3] * x[1] - c[2] * x[2],
                    -c[3] * x[0] + c[0] * x[1] + c[1] * x[2],
                     c[2] * x[0] - c[1] * x[1] + c[0] * x[2]);
    }
    Vec3 Bt(const Quat& y) const {
        return Vec3(-c[1] * y[0] + c[0] * y[1] + c[3] * y[2] - c[2] * y[3],
                    -c[2] * y[0] - c[3] * y[1] + c[0] * y[2] + c[1] * y[3],
                    -c[3] * y[0] + c[2] * y[1] - c[1] * y[2] + c[0] * y[3]);
    }
    Vec3 Ct(const Quat& y) const {
        return Vec3(-c[1] * y[0] + c[0] * y[1] - c[3] * y[2] + c[2] * y[3],
                    -c[2] * y[0] + c[3] * y[1] + c[0] * y[2] - c[1] * y[3],
                    -c[3] * y[0] - c[2] * y[1] + c[1] * y[2] + c[0] * y[3]);
    }
    /* permutations P_k q = q (x) e_k */
    Quat P(int k) const {
        if (k == 1) return Quat(-c[1], c[0], c[3], -c[2]);
        if (k == 2) return Quat(-c[2], -c[3], c[0], c[1]);
        return Quat(-c[3], c[2], -c[1], c[0]);
    }
    Vec3 A(const Vec3& x) const { return Bt(C(x)); }   /* space -> body */
    Vec3 At(const Vec3& x) const { return Ct(B(x)); }  /* body -> space */
};

/* Shepperd (1978) rotation matrix -> unit quaternion, branch structure of
   MatVec.cpp:344-373 (first maximum wins). */
Quat quatFromMatrix(const Mat3& A) {
    double a11 = A[0][0], a22 = A[1][1], a33 = A[2][2];
    double t[4] = {1.0 + a11 + a22 + a33, 1.0 + a11 - a22 - a33,
                   1.0 - a11 + a22 - a33, 1.0 - a11 - a22 + a33};
    int imax = 0;
    for (int i = 1; i < 4; i++)
        if (t[i] > t[imax]) imax = i;
    double f = 0.5 / std::sqrt(t[imax]);
    Quat q;
    switch (imax) {
    case 0:
        q[1] = (A[1][2] - A[2][1]) * f; q[2] = (A[2][0] - A[0][2]) * f; q[3] = (A[0][1] - A[1][0]) * f; break;
    case 1:
        q[0] = (A[1][2] - A[2][1]) * f; q[2] = (A[0][1] + A[1][0]) * f; q[3] = (A[0][2] + A[2][0]) * f; break;
    case 2:
        q[0] = (A[2][0] - A[0][2]) * f; q[1] = (A[0][1] + A[1][0]) * f; q[3] = (A[1][2] + A[2][1]) * f; break;
    default:
        q[0] = (A[0][1] - A[1][0]) * f; q[1] = (A[0][2] + A[2][0]) * f; q[2] = (A[1][2] + A[2][1]) * f; break;
    }
    q[imax] = t[imax] * f;
    return q;
}

/* ------------------------------------------------------------------------ */
/* Symmetric 3x3 eigen-decomposition: Smith (1961) closed-form eigenvalues,
   descending; Kopp (2008) cross-product eigenvectors.  Follows the steps of
   eigenDecomposition.cpp:37-176.                                            */
/* ------------------------------------------------------------------------ */
Vec3 eigenvalues(const Mat3& A) {
    double p1 = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    if (p1 < kEps) {
        Vec3 w = A.diag();
        if (w[0] < w[1]) std::swap(w[0], w[1]);
        if (w[0] < w[2]) std::swap(w[0], w[2]);
        if (w[1] < w[2]) std::swap(w[1], w[2]);
        return w;
    }
    Vec3 d = A.diag();
    double trace = d[0] + d[1] + d[2];
    double q = trace / 3.0;
    d -= Vec3(q, q, q);
    double p2 = d.dot(d) + 2.0 * p1;
    double p = std::sqrt(p2 / 6.0);
    double r = A.minusDiag(q).det() * (3.0 / (p * p2));
    double phi = r <= -1.0 ? kPi / 3.0 : (r >= 1.0 ? 0.0 : std::acos(r) / 3.0);
    double w0 = q + 2.0 * p * std::cos(phi);
    double w2 = q + 2.0 * p * std::cos(phi + 2.0 * kPi / 3.0);
    return Vec3(w0, trace - (w0 + w2), w2);
}

/* normalise a cross-product eigenvector candidate or fall back
   (eigenDecomposition.cpp:37-68) */
void finishEigenvector(Vec3& v, const Mat3& a, double n1base, double n2base, double thresh) {
    double norm = v.dot(v);
    double n1 = n1base + a[0][0] * a[0][0];
    double n2 = n2base + a[1][1] * a[1][1];
    double error = n1 * n2;
    if (n1 <= thresh) {
        v = Vec3(1.0, 0.0, 0.0);
    } else if (n2 <= thresh) {
        v = Vec3(0.0, 1.0, 0.0);
    } else if (norm < 4096.0 * kEps * kEps * error) {
        double t = std::fabs(a[0][1]);
        double f = -a[0][0] / a[0][1];
        if (std::fabs(a[1][1]) > t) { t = std::fabs(a[1][1]); f = -a[0][1] / a[1][1]; }
        if (std::fabs(a[1][2]) > t) f = -a[0][2] / a[1][2];
        double s = 1.0 / std::sqrt(1.0 + f * f);
        v = Vec3(s, f * s, 0.0);
    } else {
        v *= std::sqrt(1.0 / norm);
    }
}

/* rows of the returned matrix are the eigenvectors for w[0], w[1], w[2] */
Mat3 eigenvectors(const Mat3& A, const Vec3& w) {
    double wmax8eps = 8.0 * kEps * std::fabs(w[0]);
    double thresh = wmax8eps * wmax8eps;
    Mat3 a = A.symmetric();
    double n1 = a[0][1] * a[0][1] + a[0][2] * a[0][2];
    double n2 = a[0][1] * a[0][1] + a[1][2] * a[1][2];
    double base0 = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    double base1 = a[0][2] * a[0][1] - a[1][2] * a[0][0];
    double a01sq = a[0][1] * a[0][1];

    /* v0 = (A - w0) e1 x (A - w0) e2 */
    a[0][0] -= w[0];
    a[1][1] -= w[0];
    Vec3 v0(base0 + a[0][2] * w[0], base1 + a[1][2] * w[0], a[0][0] * a[1][1] - a01sq);
    finishEigenvector(v0, a, n1, n2, thresh);

    Vec3 v1;
    double t = w[0] - w[1];
    if (std::fabs(t) > wmax8eps) {
        a[0][0] += t;
        a[1][1] += t;
        v1 = Vec3(base0 + a[0][2] * w[1], base1 + a[1][2] * w[1], a[0][0] * a[1][1] - a01sq);
        finishEigenvector(v1, a, n1, n2, thresh);
    } else {
        /* degenerate w0 == w1: v1 = v0 x (A - w1) e_i for the first usable i */
        a[0][0] += w[0];
        a[1][1] += w[0];
        bool ok = false;
        for (int i = 0; i < 3 && !ok; i++) {
            a[i][i] -= w[1];
            Vec3 ai = a.col(i);
            double ni = ai.dot(ai);
            ok = ni > thresh;
            if (ok) {
                v1 = v0.cross(ai);
                double nn = v1.dot(v1);
                ok = nn > 65536.0 * kEps * kEps * ni;
                if (ok) v1 *= std::sqrt(1.0 / nn);
            }
        }
        if (!ok) {
            /* Fully degenerate: any unit vector orthogonal to v0.  The
               reference's branch here (eigenDecomposition.cpp:162-171)
               indexes q1[i+1] out of range; this restatement picks the
               normalised (v0[j], -v0[i], 0)-type vector instead
               (deviation, parity unpinned). */
            int i = 0;
            while (i < 3 && v0[i] == 0.0) i++;
            int j = (i + 1) % 3;
            double s = 1.0 / std::sqrt(v0[i] * v0[i] + v0[j] * v0[j]);
            v1 = Vec3();
            v1[i] = v0[j] * s;
            v1[j] = -v0[i] * s;
        }
    }
    return Mat3(v0, v1, v0.cross(v1));
}

/* ------------------------------------------------------------------------ */
/* Elliptic integrals and functions, written from the published methods.    */
/* ------------------------------------------------------------------------ */
const double kCarlsonTol = 1.0e-16;

/* Carlson (1995) Algorithm 2: R_C(x, y), y > 0 */
double carlsonRC(double x, double y) {
    if (!(x >= 0.0) || !(y > 0.0)) return std::nan("");
    double A0 = (x + 2.0 * y) / 3.0;
    double Q = std::pow(3.0 * kCarlsonTol, -1.0 / 8.0) * std::fabs(A0 - x);
    double xm = x, ym = y, Am = A0, scale = 1.0;
    for (int m = 0; m < 100 && scale * Q >= std::fabs(Am); m++) {
        double lam = 2.0 * std::sqrt(xm) * std::sqrt(ym) + ym;
        xm = 0.25 * (xm + lam);
        ym = 0.25 * (ym + lam);
        Am = 0.25 * (Am + lam);
        scale *= 0.25;
    }
    double s = (y - A0) * scale / Am;
    double poly = 1.0 + s * s * (3.0 / 10.0 + s * (1.0 / 7.0 + s * (3.0 / 8.0 + s * (9.0 / 22.0 +
                  s * (159.0 / 208.0 + s * (9.0 / 8.0))))));
    return poly / std::sqrt(Am);
}

/* Carlson (1995) Algorithm 1: R_F(x, y, z), at most one argument zero */
double carlsonRF(double x, double y, double z) {
    if (!(x >= 0.0 && y >= 0.0 && z >= 0.0) || x + y == 0.0 || x + z == 0.0 || y + z == 0.0)
        return std::nan("");
    double A0 = (x + y + z) / 3.0;
    double Q = std::pow(3.0 * kCarlsonTol, -1.0 / 6.0) *
               std::max(std::fabs(A0 - x), std::max(std::fabs(A0 - y), std::fabs(A0 - z)));
    double xm = x, ym = y, zm = z, Am = A0, scale = 1.0;
    for (int m = 0; m < 100 && scale * Q >= std::fabs(Am); m++) {
        double sx = std::sqrt(xm), sy = std::sqrt(ym), sz = std::sqrt(zm);
        double lam = sx * sy + sx * sz + sy * sz;
        xm = 0.25 * (xm + lam); ym = 0.25 * (ym + lam); zm = 0.25 * (zm + lam);
        Am = 0.25 * (Am + lam);
        scale *= 0.25;
    }
    double X = (A0 - x) * scale / Am;
    double Y = (A0 - y) * scale / Am;
    double Z = -(X + Y);
    double E2 = X * Y - Z * Z;
    double E3 = X * Y * Z;
    double poly = 1.0 - E2 / 10.0 + E3 / 14.0 + E2 * E2 / 24.0 - 3.0 * E2 * E3 / 44.0;
    return poly / std::sqrt(Am);
}

/* Carlson (1995) Algorithm 3: R_J(x, y, z, p), p > 0 */
double carlsonRJ(double x, double y, double z, double p) {
    if (!(x >= 0.0 && y >= 0.0 && z >= 0.0 && p > 0.0) || x + y == 0.0 || x + z == 0.0 || y + z == 0.0)
        return std::nan("");
    double A0 = (x + y + z + 2.0 * p) / 5.0;
    double delta = (p - x) * (p - y) * (p - z);
    double Q = std::pow(0.25 * kCarlsonTol, -1.0 / 6.0) *
               std::max(std::max(std::fabs(A0 - x), std::fabs(A0 - y)),
                        std::max(std::fabs(A0 - z), std::fabs(A0 - p)));
    double xm = x, ym = y, zm = z, pm = p, Am = A0, scale = 1.0, scale3 = 1.0;
    double sum = 0.0;
    for (int m = 0; m < 100 && scale * Q >= std::fabs(Am); m++) {
        double sx = std::sqrt(xm), sy = std::sqrt(ym), sz = std::sqrt(zm), sp = std::sqrt(pm);
        double lam = sx * sy + sx * sz + sy * sz;
        double d = (sp + sx) * (sp + sy) * (sp + sz);
        double e = scale3 * delta / (d * d);
        sum += scale * carlsonRC(1.0, 1.0 + e) / d;
        xm = 0.25 * (xm + lam); ym = 0.25 * (ym + lam); zm = 0.25 * (zm + lam); pm = 0.25 * (pm + lam);
        Am = 0.25 * (Am + lam);
        scale *= 0.25;
        scale3 *= 1.0 / 64.0;
    }
    double X = (A0 - x) * scale / Am;
    double Y = (A0 - y) * scale / Am;
    double Z = (A0 - z) * scale / Am;
    double P = -0.5 * (X + Y + Z);
    double E2 = X * Y + X * Z + Y * Z - 3.0 * P * P;
    double E3 = X * Y * Z + 2.0 * E2 * P + 4.0 * P * P * P;
    double E4 = (2.0 * X * Y * Z + E2 * P + 3.0 * P * P * P) * P;
    double E5 = X * Y * Z * P * P;
    double poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 -
                  9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return scale * poly / (Am * std::sqrt(Am)) + 6.0 * sum;
}

/* Omega(x, n, m) = -(n x^3 / 3) R_J(1 - x^2, 1 - m x^2, 1, 1 + n x^2), the
   function exactRotation uses (definition: ellipticFunctions.h:251-254) */
double omegaFn(double x, double n, double m) {
    double x2 = x * x;
    return (-1.0 / 3.0) * n * x * x2 * carlsonRJ(1.0 - x2, 1.0 - m * x2, 1.0, 1.0 + n * x2);
}

/* Jacobi sn, cn, dn(u | m), 0 <= m <= 1, by the arithmetic-geometric mean and
   descending Landen transformation (DLMF 22.20(ii)). */
void jacobiFn(double u, double m, double& sn, double& cn, double& dn) {
    if (std::fabs(m) > 1.0 || m < 0.0) { sn = cn = dn = std::nan(""); return; }
    if (m < 2.0 * kEps) { sn = std::sin(u); cn = std::cos(u); dn = 1.0; return; }
    if (1.0 - m < 2.0 * kEps) { sn = std::tanh(u); cn = 1.0 / std::cosh(u); dn = cn; return; }
    /* sn and cn have real period 4K, dn 2K: reduce u to [-2K, 2K] so the
       scaled phase 2^N a_N u below stays small (DLMF 22.4(ii)) */
    double K = carlsonRF(0.0, 1.0 - m, 1.0);
    if (std::fabs(u) > 2.0 * K) u -= 4.0 * K * std::nearbyint(u / (4.0 * K));
    const int kMax = 24;
    double a[kMax + 1], c[kMax + 1];
    a[0] = 1.0;
    c[0] = std::sqrt(m);
    double b = std::sqrt(1.0 - m);
    int n = 0;
    while (std::fabs(c[n]) > kEps * a[n] && n < kMax) {
        a[n + 1] = 0.5 * (a[n] + b);
        c[n + 1] = 0.5 * (a[n] - b);
        b = std::sqrt(a[n] * b);
        n++;
    }
    double phi = std::ldexp(a[n] * u, n);
    for (int i = n; i > 0; i--) phi = 0.5 * (phi + std::asin(c[i] / a[i] * std::sin(phi)));
    sn = std::sin(phi);
    cn = std::cos(phi);
    /* dn^2 = (1 - m) + m cn^2 (DLMF 22.6.1): both terms are non-negative, so
       this stays accurate near u = K where cn -> 0 */
    dn = std::sqrt((1.0 - m) + m * cn * cn);
}

/* ------------------------------------------------------------------------ */
/* RigidBody (RigidBody.h:18-53)                                            */
/* ------------------------------------------------------------------------ */
struct Body {
    int N = 0;
    int dof = 0;
    double mass = 0.0;
    Vec3 I, rcm, pcm;
    Quat q, pi;
    Vec3 force;
    Quat torque;
    int loc = 0;
    double invMass = 0.0;
    Vec3 invI;
    double twoKt = 0.0, twoKr = 0.0;
    std::vector<Vec3> delta;
};

/* collinearity test, RigidBody.cpp:26-45 */
bool collinear(const std::vector<Vec3>& delta, const std::vector<double>& d2, Vec3& u) {
    const double TOL = 1.0e-5;
    double d0 = d2[0];
    int jmax = 0;
    double d2max = d0;
    int N = (int)delta.size();
    for (int j = 1; j < N; j++)
        if (d2[j] > d0) { jmax = j; d2max = d2[j]; }
    u = delta[jmax] / std::sqrt(d2max);
    bool all = true;
    for (int j = 0; all && j < N; j++) {
        double udj = u.dot(delta[j]);
        all = d2[j] < TOL * d2max || std::fabs(udj * udj / d2[j] - 1.0) < TOL;
    }
    return all;
}

/* unit vector orthogonal to u, RigidBody.cpp:51-58: Projection(u) e_m
   normalised, m = index of the smallest signed component of u.  For u = -e_k
   that e_m is parallel to u and the result is 0/0 (NaN); with
   RBREF_FIX_ORTHONORMAL m is the index of the smallest |u_k|, which keeps
   |Projection(u) e_m| >= sqrt(2/3). */
Vec3 orthonormal(const Vec3& u, unsigned flags) {
    int imin;
    if (flags & RBREF_FIX_ORTHONORMAL) {
        imin = std::fabs(u[0]) < std::fabs(u[1]) ? 0 : 1;
        if (std::fabs(u[2]) < std::fabs(u[imin])) imin = 2;
    } else {
        imin = u[0] < u[1] ? 0 : 1;
        if (u[2] < u[imin]) imin = 2;
    }
    Vec3 v;
    v[imin] = 1.0;
    v = projection(u) * v;
    return v / std::sqrt(v.dot(v));
}

/* uniaxial rotation about principal axis k, RigidBody.cpp:189-214 */
void uniaxial(Quat& q, Quat& pi, const Vec3& invI, int k, double dt) {
    Quat Bq = q.P(k);
    double phi = 0.25 * pi.dot(Bq) * dt * invI[k - 1];
    double s = std::sin(phi), c = std::cos(phi);
    Quat Bpi = pi.P(k);
    q = q * c + Bq * s;
    pi = pi * c + Bpi * s;
}

/* NO-SQUISH, RigidBody.cpp:220-231 */
void noSquish(Quat& q, Quat& pi, const Vec3& invI, int dof, double dt, int n) {
    double h = dt / n;
    double h2 = 0.5 * h;
    bool axis3 = dof == 6;
    for (int i = 0; i < n; i++) {
        if (axis3) uniaxial(q, pi, invI, 3, h2);
        uniaxial(q, pi, invI, 2, h2);
        uniaxial(q, pi, invI, 1, h);
        uniaxial(q, pi, invI, 2, h2);
        if (axis3) uniaxial(q, pi, invI, 3, h2);
    }
}

inline int signOf(double x) { return x >= 0 ? 1 : -1; }
/* round half toward zero, as stairCase in RigidBody.cpp:238 */
inline int stair(double x) { return x > 0 ? (int)std::ceil(x - 0.5) : (int)std::floor(x + 0.5); }

/* closed-form free linear rotor (policy for dof-5 bodies in exact mode):
   body-frame L is constant and the body turns about it at |w| = |L|/I0 */
void linearRotorExact(Quat& q, Quat& pi, const Vec3& invI, double dt) {
    Vec3 L = q.Bt(pi) * 0.5;
    Vec3 w = diagMul(invI, L);
    double wn = std::sqrt(w.dot(w));
    if (wn == 0.0) return;
    double half = 0.5 * wn * dt;
    Vec3 axis = w / wn;
    q = q * std::cos(half) + q.B(axis) * std::sin(half);
    q /= q.norm();
    pi = q.B(L * 2.0);
}

/* exact free-rotor rotation, RigidBody.cpp:241-308 (same sequence of
   operations; elliptic functions from the routines above) */
void exactRotation(Quat& q, Quat& pi, const Vec3& I, const Vec3& invI, double dt) {
    Vec3 Iw = q.Bt(pi) * 0.5;
    Vec3 w0 = diagMul(invI, Iw);
    double Lsq = Iw[1] * Iw[1] + Iw[2] * Iw[2];
    if (Lsq < kEps) { uniaxial(q, pi, invI, 1, dt); return; }
    Lsq += Iw[0] * Iw[0];
    double L = std::sqrt(Lsq);
    double twoKr = Iw.dot(w0);
    Quat z0(Iw[2], Iw[1], L - Iw[0], 0.0);
    double r1 = Lsq - twoKr * I[2];
    double r3 = twoKr * I[0] - Lsq;
    double l1 = r1 * invI[1] / (I[1] - I[2]);
    double l3 = r3 * invI[1] / (I[0] - I[1]);
    double lmin = std::min(l1, l3);
    double c13 = 1.0 / (I[0] - I[2]);
    Vec3 a(signOf(w0[0]) * std::sqrt(r1 * invI[0] * c13), std::sqrt(lmin),
           signOf(w0[2]) * std::sqrt(r3 * invI[2] * c13));
    double m = lmin / std::max(l1, l3);
    double K = carlsonRF(0.0, 1.0 - m, 1.0);
    double inv2K = 0.5 / K;
    double s0 = w0[1] / a[1];
    double c0, u0;
    int i0;
    if (std::fabs(s0) < 1.0) {
        c0 = l1 < l3 ? w0[0] / a[0] : w0[2] / a[2];
        u0 = s0 * carlsonRF(1.0 - s0 * s0, 1.0 - m * s0 * s0, 1.0);
        i0 = stair(u0 * inv2K);
    } else {
        a[1] = std::fabs(w0[1]);
        s0 = signOf(s0);
        c0 = 0.0;
        u0 = s0 * K;
        i0 = 0;
    }
    double wp = -invI[1] * a[0] * a[2] / (a[1] * c13);
    double u = wp * dt + u0;
    int jump = stair(u * inv2K) - i0;
    double sn, cn, dn, deltaF;
    jacobiFn(u, m, sn, cn, dn);
    double alpha = I[0] * a[0] / L;
    double eta = alpha * alpha;
    eta /= 1.0 - eta;
    Vec3 Ia = diagMul(I, a);
    if (l1 < l3) {
        double C = std::sqrt(m + eta);
        deltaF = u - u0 + signOf(cn) * omegaFn(sn, eta, m) - signOf(c0) * omegaFn(s0, eta, m) +
                 (alpha / C) * (std::atan(C * sn / dn) - std::atan(C * s0 * a[2] / w0[2]));
        if (jump != 0) deltaF += jump * 2.0 * omegaFn(1.0, eta, m);
        Iw = diagMul(Ia, Vec3(cn, sn, dn));
    } else {
        double k2eta = m * eta;
        double C = std::sqrt(1.0 + k2eta);
        deltaF = u - u0 + signOf(cn) * omegaFn(sn, k2eta, m) - signOf(c0) * omegaFn(s0, k2eta, m) +
                 (alpha / C) * (std::atan(C * sn / cn) - std::atan(C * s0 / c0));
        if (jump != 0) deltaF += jump * (2.0 * omegaFn(1.0, k2eta, m) + (alpha / C) * kPi);
        Iw = diagMul(Ia, Vec3(dn, sn, cn));
    }
    deltaF *= 1.0 + eta;
    double theta = (Lsq * (u - u0) + r3 * deltaF) / (2.0 * L * I[0] * wp);
    Quat z = Quat(Iw[2], Iw[1], L - Iw[0], 0.0) * std::cos(theta) +
             Quat(-Iw[1], Iw[2], 0.0, L - Iw[0]) * std::sin(theta);
    q = z * z0.dot(q) + z.C(z0.Ct(q));
    q /= q.norm();
    pi = q.B(Iw * 2.0);
}

void rotateBody(Quat& q, Quat& pi, const Vec3& I, const Vec3& invI, int dof, double dt, int mode,
                unsigned flags) {
    if (mode == 0) {
        if (dof == 5 && (flags & RBREF_FIX_LINEAR_EXACT))
            linearRotorExact(q, pi, invI, dt);
        else
            exactRotation(q, pi, I, invI, dt);
    } else {
        noSquish(q, pi, invI, dof, dt, mode);
    }
}

/* ------------------------------------------------------------------------ */
/* cleanBodyIndices, RigidBodySystem.cpp:28-49: distinct positive labels are
   renumbered 1..B in ascending label order; labels <= 0 become 0.           */
/* ------------------------------------------------------------------------ */
std::vector<int> cleanBodyIndices(const std::vector<int>& in) {
    int n = (int)in.size();
    int maxIndex = 0;
    for (int v : in) maxIndex = std::max(maxIndex, v);
    std::vector<int> first(maxIndex > 0 ? maxIndex : 0, -1), next(n, -1), out(n, 0);
    for (int i = 0; i < n; i++)
        if (in[i] > 0) { next[i] = first[in[i] - 1]; first[in[i] - 1] = i; }
    int body = 0;
    for (int label = 0; label < maxIndex; label++) {
        if (first[label] == -1) continue;
        body++;
        for (int j = first[label]; j != -1; j = next[j]) out[j] = body;
    }
    return out;
}

/* ------------------------------------------------------------------------ */
/* Context stand-in: FP64 positions/velocities/forces + the force field.    */
/* ------------------------------------------------------------------------ */
struct Bond { int a, b; double length, k; };

struct VirtualSite { int type; int site; int p[3]; double w[12]; };

struct SystemState {
    int n = 0;
    std::vector<double> mass;
    std::vector<Vec3> R, V, F;
    std::vector<double> tetherK;
    std::vector<Vec3> tetherR0;
    std::vector<Bond> bonds;
    std::vector<VirtualSite> sites;
    double pe = 0.0;
    double time = 0.0;
    long stepCount = 0;

    /* calcForcesAndEnergy(true, false): tether + harmonic bonds */
    void calcForces() {
        pe = 0.0;
        for (int i = 0; i < n; i++) F[i] = Vec3();
        if (!tetherK.empty())
            for (int i = 0; i < n; i++) {
                Vec3 dr = R[i] - tetherR0[i];
                F[i] = dr * (-tetherK[i]);
                pe += 0.5 * tetherK[i] * dr.dot(dr);
            }
        for (const Bond& b : bonds) {
            Vec3 dr = R[b.b] - R[b.a];
            double r = std::sqrt(dr.dot(dr));
            double dEdR = b.k * (r - b.length);
            pe += 0.5 * b.k * (r - b.length) * (r - b.length);
            Vec3 f = dr * (dEdR / r);
            F[b.a] += f;
            F[b.b] -= f;
        }
        distributeSiteForces();
    }

    /* OpenMM's virtual sites (ReferenceVirtualSites, called at
       ReferenceRigidBodyKernels.cpp:100), written from their published
       definitions: TwoParticleAverage r = w1 r1 + w2 r2, ThreeParticleAverage
       r = w1 r1 + w2 r2 + w3 r3, OutOfPlane r = r1 + w12 r12 + w13 r13 +
       wc r12 x r13 (r1j = rj - r1), LocalCoordinates r = o + p_x e_x +
       p_y e_y + p_z e_z with o, x, y the origin-, x- and y-weighted sums of
       the parents, e_x = x/|x|, e_z = (x × y)/|x × y|, e_y = e_z × e_x
       (weights: origin 0-2, x 3-5, y 6-8, local position 9-11) */
    struct LocalFrame { Vec3 x, y, ex, ey, ez; double nx, nz; };
    LocalFrame localFrame(const VirtualSite& v) const {
        LocalFrame f;
        for (int k = 0; k < 3; k++) {
            f.x += R[v.p[k]] * v.w[3 + k];
            f.y += R[v.p[k]] * v.w[6 + k];
        }
        Vec3 z = f.x.cross(f.y);
        f.nx = std::sqrt(f.x.dot(f.x));
        f.nz = std::sqrt(z.dot(z));
        f.ex = f.x / f.nx;
        f.ez = z / f.nz;
        f.ey = f.ez.cross(f.ex);
        return f;
    }
    void computeSitePositions() {
        for (const VirtualSite& v : sites) {
            const Vec3 &r1 = R[v.p[0]], &r2 = R[v.p[1]];
            if (v.type == RBREF_SITE_LOCAL_COORDINATES) {
                LocalFrame f = localFrame(v);
                Vec3 o = R[v.p[0]] * v.w[0] + R[v.p[1]] * v.w[1] + R[v.p[2]] * v.w[2];
                R[v.site] = o + f.ex * v.w[9] + f.ey * v.w[10] + f.ez * v.w[11];
            } else if (v.type == RBREF_SITE_TWO_AVERAGE) {
                R[v.site] = r1 * v.w[0] + r2 * v.w[1];
            } else if (v.type == RBREF_SITE_THREE_AVERAGE) {
                R[v.site] = r1 * v.w[0] + r2 * v.w[1] + R[v.p[2]] * v.w[2];
            } else {
                Vec3 r12 = r2 - r1, r13 = R[v.p[2]] - r1;
                R[v.site] = r1 + r12 * v.w[0] + r13 * v.w[1] + r12.cross(r13) * v.w[2];
            }
        }
    }
    /* the force on a site moved to its parents by the chain rule (the
       transpose of d r_site / d r_parent), the site's own force cleared */
    void distributeSiteForces() {
        for (const VirtualSite& v : sites) {
            Vec3 f = F[v.site];
            F[v.site] = Vec3();
            if (v.type == RBREF_SITE_TWO_AVERAGE) {
                F[v.p[0]] += f * v.w[0];
                F[v.p[1]] += f * v.w[1];
            } else if (v.type == RBREF_SITE_THREE_AVERAGE) {
                F[v.p[0]] += f * v.w[0];
                F[v.p[1]] += f * v.w[1];
                F[v.p[2]] += f * v.w[2];
            } else if (v.type == RBREF_SITE_LOCAL_COORDINATES) {
                /* F_i = ow_i F + xw_i Gx + yw_i Gy with Gx, Gy the transposed
                   derivatives of r with respect to x and y: Jx, Jz project
                   out e_x, e_z and divide by |x|, |x × y| */
                LocalFrame fr = localFrame(v);
                auto Jx = [&](Vec3 a) { return (a - fr.ex * fr.ex.dot(a)) / fr.nx; };
                auto Jz = [&](Vec3 a) { return (a - fr.ez * fr.ez.dot(a)) / fr.nz; };
                const double px = v.w[9], py = v.w[10], pz = v.w[11];
                Vec3 jzexf = Jz(fr.ex.cross(f)), jzf = Jz(f);
                Vec3 gx = Jx(f) * px + (fr.y.cross(jzexf) - Jx(fr.ez.cross(f))) * py + fr.y.cross(jzf) * pz;
                Vec3 gy = fr.x.cross(jzexf) * (-py) - fr.x.cross(jzf) * pz;
                for (int k = 0; k < 3; k++) F[v.p[k]] += f * v.w[k] + gx * v.w[3 + k] + gy * v.w[6 + k];
            } else {
                Vec3 r12 = R[v.p[1]] - R[v.p[0]], r13 = R[v.p[2]] - R[v.p[0]];
                Vec3 f2 = f * v.w[0] + r13.cross(f) * v.w[2];
                Vec3 f3 = f * v.w[1] + f.cross(r12) * v.w[2];
                F[v.p[0]] += f - f2 - f3;
                F[v.p[1]] += f2;
                F[v.p[2]] += f3;
            }
        }
    }
};

/* ------------------------------------------------------------------------ */
/* RigidBodySystem, RigidBodySystem.h:27-61 / RigidBodySystem.cpp            */
/* ------------------------------------------------------------------------ */
struct FreeAtom { int index; double invMass; Vec3 savedPos; };

struct BodySystem {
    int rotationMode = 0;
    unsigned flags = 0;
    std::vector<int> bodyIndex, atomIndex;
    std::vector<Body> body;
    std::vector<Vec3> d;          /* bodyFixedPositions */
    std::vector<FreeAtom> freeAtom;
    int numFree = 0, numBodies = 0, numActualAtoms = 0, numBodyAtoms = 0, numDOF = 0;
    int numConstraints = 0;
    double transKE = 0.0, rotKE = 0.0;

    int atomOf(const Body& b, int j) const { return atomIndex[numFree + b.loc + j]; }

    /* RigidBodySystem::initialize, RigidBodySystem.cpp:55-114 */
    void initialize(const SystemState& ctx, const std::vector<int>& bodyIndices, int mode, unsigned fl) {
        bodyIndex = cleanBodyIndices(bodyIndices);
        rotationMode = mode;
        flags = fl;
        numBodies = 0;
        for (int v : bodyIndex) numBodies = std::max(numBodies, v);
        int numAtoms = ctx.n;
        numActualAtoms = numAtoms;   /* no virtual sites in this system model */
        atomIndex.assign(numActualAtoms, 0);
        numFree = 0;
        body.assign(numBodies, Body());
        /* massless particles (virtual sites are massless) are neither free
           nor body atoms (RigidBodySystem.cpp:66-76); they must be labelled
           free, because the reference's body-atom loop (:99-104) does not
           skip them and would overwrite the next body's entries */
        for (int i = 0; i < numAtoms; i++) {
            if (ctx.mass[i] == 0.0) {
                if (bodyIndex[i] != 0)
                    throw std::runtime_error("massless particles cannot belong to rigid bodies");
                continue;
            }
            if (bodyIndex[i] == 0) atomIndex[numFree++] = i;
            else body[bodyIndex[i] - 1].N++;
        }
        numBodyAtoms = numActualAtoms - numFree;
        d.assign(numBodyAtoms, Vec3());
        if (numActualAtoms < 0) throw std::runtime_error("internal: atom counts");
        freeAtom.resize(numFree);
        for (int k = 0; k < numFree; k++) {
            freeAtom[k].index = atomIndex[k];
            freeAtom[k].invMass = 1.0 / ctx.mass[freeAtom[k].index];
        }
        int loc = 0;
        for (Body& b : body) {
            if (b.N == 1)
                throw std::runtime_error("a rigid body must contain at least two atoms");
            b.loc = loc;
            loc += b.N;
        }
        std::vector<int> fill(numBodies, 0);
        for (int i = 0; i < numAtoms; i++) {
            int ib = bodyIndex[i];
            if (ib > 0) {
                Body& b = body[ib - 1];
                atomIndex[numFree + b.loc + fill[ib - 1]++] = i;
            }
        }
    }

    /* RigidBody::forceAndTorque, RigidBody.cpp:174-183 */
    void forceAndTorque(Body& b, const std::vector<Vec3>& F) const {
        b.force = Vec3();
        Vec3 tau;
        for (int j = 0; j < b.N; j++) {
            int i = atomOf(b, j);
            b.force += F[i];
            tau += b.delta[j].cross(F[i]);
        }
        b.torque = b.q.C(tau);
    }

    /* RigidBody::buildGeometry, RigidBody.cpp:65-116 */
    void buildGeometry(Body& b, const std::vector<Vec3>& R, const std::vector<Vec3>& F,
                       const std::vector<double>& M) {
        b.mass = 0.0;
        b.rcm = Vec3();
        for (int j = 0; j < b.N; j++) {
            int i = atomOf(b, j);
            b.mass += M[i];
            b.rcm += R[i] * M[i];
        }
        b.rcm /= b.mass;
        b.invMass = 1.0 / b.mass;
        b.delta.resize(b.N);
        std::vector<double> d2(b.N);
        for (int j = 0; j < b.N; j++) {
            b.delta[j] = R[atomOf(b, j)] - b.rcm;
            d2[j] = b.delta[j].dot(b.delta[j]);
        }
        Vec3 u;
        Mat3 A;
        if (collinear(b.delta, d2, u)) {
            double moi = 0.0;
            for (int j = 0; j < b.N; j++) moi += M[atomOf(b, j)] * d2[j];
            b.I = Vec3(moi, moi, 0.0);
            b.invI = Vec3(1.0 / moi, 1.0 / moi, 0.0);
            Vec3 v = orthonormal(u, flags);
            Mat3 rows(v, u.cross(v), u);
            A = (flags & RBREF_FIX_LINEAR_FRAME) ? rows : rows.t();
            b.dof = 5;
        } else {
            Mat3 inertia;
            for (int j = 0; j < b.N; j++) inertia += projection(b.delta[j]) * M[atomOf(b, j)];
            b.I = eigenvalues(inertia);
            b.invI = Vec3(1.0 / b.I[0], 1.0 / b.I[1], 1.0 / b.I[2]);
            A = eigenvectors(inertia, b.I);
            b.dof = 6;
        }
        b.q = quatFromMatrix(A);
        for (int j = 0; j < b.N; j++) d[b.loc + j] = A * b.delta[j];
        forceAndTorque(b, F);
    }

    /* RigidBody::buildDynamics, RigidBody.cpp:123-142 */
    void buildDynamics(Body& b, const std::vector<Vec3>& V, const std::vector<double>& M) {
        if (flags & RBREF_FIX_PCM_RESET) b.pcm = Vec3();
        for (int j = 0; j < b.N; j++) {
            int i = atomOf(b, j);
            b.pcm += V[i] * M[i];
        }
        Vec3 vcm = b.pcm / b.mass;
        b.twoKt = b.pcm.dot(vcm);
        Vec3 L;
        for (int j = 0; j < b.N; j++) {
            int i = atomOf(b, j);
            L += d[b.loc + j].cross(b.q.A(V[i] - vcm) * M[i]);
        }
        b.pi = b.q.B(L) * 2.0;
        b.twoKr = L.dot(diagMul(b.invI, L));
    }

    /* RigidBodySystem::update, RigidBodySystem.cpp:120-142 */
    void update(const SystemState& ctx, bool geometry, bool velocities) {
        if (geometry) {
            numDOF = numFree - numConstraints;
            for (Body& b : body) { buildGeometry(b, ctx.R, ctx.F, ctx.mass); numDOF += b.dof; }
        }
        if (velocities)
            for (Body& b : body) buildDynamics(b, ctx.V, ctx.mass);
    }

    /* RigidBody::updateAtomicPositions, RigidBody.cpp:148-153 */
    void updatePositions(Body& b, std::vector<Vec3>& R) {
        for (int j = 0; j < b.N; j++) {
            b.delta[j] = b.q.At(d[b.loc + j]);
            R[atomOf(b, j)] = b.rcm + b.delta[j];
        }
    }

    /* RigidBody::updateAtomicVelocities, RigidBody.cpp:159-168 */
    void updateVelocities(Body& b, std::vector<Vec3>& V) {
        Vec3 L = b.q.Bt(b.pi) * 0.5;
        Vec3 omega = diagMul(b.invI, L);
        Vec3 ws = b.q.At(omega);
        Vec3 vcm = b.pcm * b.invMass;
        for (int j = 0; j < b.N; j++)
            V[atomOf(b, j)] = vcm + ((flags & RBREF_PLANT_VELOCITY_SIGN) ? b.delta[j].cross(ws) : ws.cross(b.delta[j]));
        b.twoKt = b.pcm.dot(vcm);
        b.twoKr = L.dot(omega);
    }

    /* RigidBodySystem::integratePart1, RigidBodySystem.cpp:170-187 */
    void integratePart1(double dt, const std::vector<Vec3>& F, std::vector<Vec3>& V, std::vector<Vec3>& R) {
        double halfDt = 0.5 * dt;
        for (FreeAtom& a : freeAtom) {
            V[a.index] += F[a.index] * a.invMass * halfDt;
            R[a.index] += V[a.index] * dt;
            a.savedPos = R[a.index];
        }
        for (Body& b : body) {
            b.pcm += b.force * halfDt;
            b.pi += b.torque * dt;
            b.rcm += b.pcm * (b.invMass * dt);
            rotateBody(b.q, b.pi, b.I, b.invI, b.dof, dt, rotationMode, flags);
            updatePositions(b, R);
        }
    }

    /* RigidBodySystem::integratePart2, RigidBodySystem.cpp:193-204 */
    void integratePart2(double dt, const std::vector<Vec3>& R, const std::vector<Vec3>& F, std::vector<Vec3>& V) {
        double halfDt = 0.5 * dt;
        double invDt = 1.0 / dt;
        for (FreeAtom& a : freeAtom)
            V[a.index] += F[a.index] * a.invMass * halfDt + (R[a.index] - a.savedPos) * invDt;
        for (Body& b : body) {
            forceAndTorque(b, F);
            b.pcm += b.force * halfDt;
            b.pi += b.torque * dt;
            updateVelocities(b, V);
        }
    }

    /* RigidBodySystem::computeKineticEnergies, RigidBodySystem.cpp:210-220 */
    void computeKineticEnergies(const std::vector<Vec3>& V) {
        transKE = rotKE = 0.0;
        for (const FreeAtom& a : freeAtom) transKE += V[a.index].dot(V[a.index]) / a.invMass;
        for (const Body& b : body) { transKE += b.twoKt; rotKE += b.twoKr; }
        transKE *= 0.5;
        rotKE *= 0.5;
    }
};

/* ------------------------------------------------------------------------ */
/* Distance constraints between free atoms.  The reference hands them to      */
/* OpenMM's ReferenceConstraints: positions after integratePart1, with the    */
/* positions before the step as the reference directions, and velocities     */
/* after integratePart2 (ReferenceRigidBodyKernels.cpp:92-104); constraints   */
/* touching body atoms are refused (RigidBodySystem.cpp:107-112) and each one */
/* removes a degree of freedom (RigidBodySystem.cpp:130).  OpenMM is not part */
/* of the reference tree, so its constraint solver is not restated: this is  */
/* SHAKE for positions (Ryckaert, Ciccotti and Berendsen, J. Comput. Phys.    */
/* 23, 327 (1977)) and RATTLE's velocity stage (Andersen, J. Comput. Phys.    */
/* 52, 24 (1983)), Gauss-Seidel over each connected cluster of constraints in */
/* the order they were added.  Positions are converged when every |r_ij| is   */
/* within (1 +- tol) d_ij, OpenMM's criterion; velocities when every          */
/* |r_ij . v_ij| <= tol |r_ij| |v_ij|.  The converged results are fixed by    */
/* the constraint equations, which the tests check (parity of the iteration   */
/* itself with OpenMM is unpinned).                                           */
/* ------------------------------------------------------------------------ */
struct Constraint { int a, b; double d; };

struct Constraints {
    static constexpr int kMaxIterations = 1000;
    std::vector<Constraint> list;
    std::vector<std::vector<int>> clusters;   /* constraint indices, in order added */
    double tol = 1e-5;
    long failures = 0;

    /* connected components of the constraint graph */
    void build(int n) {
        std::vector<int> root(n);
        for (int i = 0; i < n; i++) root[i] = i;
        auto find = [&](int x) { while (root[x] != x) x = root[x] = root[root[x]]; return x; };
        for (const Constraint& c : list) root[find(c.a)] = find(c.b);
        std::vector<int> id(n, -1);
        clusters.clear();
        for (size_t k = 0; k < list.size(); k++) {
            int r = find(list[k].a);
            if (id[r] < 0) { id[r] = (int)clusters.size(); clusters.emplace_back(); }
            clusters[id[r]].push_back((int)k);
        }
    }

    void applyPositions(const std::vector<Vec3>& Rold, std::vector<Vec3>& R, const std::vector<double>& invMass) {
        const double lo = (1.0 - tol) * (1.0 - tol), hi = (1.0 + tol) * (1.0 + tol);
        for (const std::vector<int>& cl : clusters) {
            bool done = false;
            for (int it = 0; it < kMaxIterations && !done; it++) {
                done = true;
                for (int k : cl) {
                    const Constraint& c = list[k];
                    Vec3 s = R[c.a] - R[c.b];
                    double s2 = s.dot(s), d2 = c.d * c.d;
                    if (s2 >= lo * d2 && s2 <= hi * d2) continue;
                    done = false;
                    Vec3 ro = Rold[c.a] - Rold[c.b];
                    double g = (d2 - s2) / (2.0 * (invMass[c.a] + invMass[c.b]) * s.dot(ro));
                    R[c.a] += ro * (g * invMass[c.a]);
                    R[c.b] -= ro * (g * invMass[c.b]);
                }
            }
            if (!done) failures++;
        }
    }

    void applyVelocities(const std::vector<Vec3>& R, std::vector<Vec3>& V, const std::vector<double>& invMass) {
        const double tol2 = tol * tol;
        for (const std::vector<int>& cl : clusters) {
            bool done = false;
            for (int it = 0; it < kMaxIterations && !done; it++) {
                done = true;
                for (int k : cl) {
                    const Constraint& c = list[k];
                    Vec3 s = R[c.a] - R[c.b];
                    Vec3 dv = V[c.a] - V[c.b];
                    double rv = s.dot(dv), s2 = s.dot(s);
                    if (rv * rv <= tol2 * s2 * dv.dot(dv)) continue;
                    done = false;
                    double g = -rv / ((invMass[c.a] + invMass[c.b]) * s2);
                    V[c.a] += s * (g * invMass[c.a]);
                    V[c.b] -= s * (g * invMass[c.b]);
                }
            }
            if (!done) failures++;
        }
    }
};

}  // namespace

/* ------------------------------------------------------------------------ */
/* Integrator + Reference kernel (RigidBodyIntegrator.cpp,
   ReferenceRigidBodyKernels.cpp)                                            */
/* ------------------------------------------------------------------------ */
struct rbref_sim {
    SystemState ctx;
    std::vector<int> bodyIndices;
    double dt = 0.0;
    int rotationMode = 0;
    unsigned flags = 0;
    BodySystem integratorSystem;   /* RigidBodyIntegrator::bodySystem */
    BodySystem kernelSystem;       /* ReferenceIntegrateRigidBodyStepKernel::bodySystem */
    std::vector<Vec3> oldPos;
    Constraints constraints;
    std::vector<double> invMass;   /* 1/m per atom, for the constraints */

    /* RigidBodyIntegrator::stateChanged, RigidBodyIntegrator.cpp:63-74, with
       uploadBodySystem = RigidBodySystem::copy (ReferenceRigidBodyKernels.cpp:78-80) */
    void stateChanged(bool positions) {
        stateChangedBodies(positions);
        refineReset();
    }
    void stateChangedBodies(bool positions) {
        if (positions) {
            ctx.calcForces();
            integratorSystem.update(ctx, true, true);
            kernelSystem = integratorSystem;
        } else if (flags & RBREF_FIX_STALE_SYSTEM) {
            kernelSystem.update(ctx, false, true);
            integratorSystem = kernelSystem;
        } else {
            integratorSystem.update(ctx, false, true);
            kernelSystem = integratorSystem;
        }
    }

    /* ReferenceIntegrateRigidBodyStepKernel::execute, ReferenceRigidBodyKernels.cpp:82-108,
       with the refined-energy accumulation of the CUDA platform around it */
    void execute() {
        if (refined) refineBegin();
        for (int k = 0; k < kernelSystem.numFree; k++) {
            int i = kernelSystem.atomIndex[k];
            oldPos[i] = ctx.R[i];
        }
        kernelSystem.integratePart1(dt, ctx.F, ctx.V, ctx.R);
        if (kernelSystem.numFree != 0 && !constraints.list.empty())
            constraints.applyPositions(oldPos, ctx.R, invMass);
        ctx.computeSitePositions();
        ctx.calcForces();
        kernelSystem.integratePart2(dt, ctx.R, ctx.F, ctx.V);
        if (kernelSystem.numFree != 0 && !constraints.list.empty())
            constraints.applyVelocities(ctx.R, ctx.V, invMass);
        if (refined) refineEnd();
        ctx.time += dt;
        ctx.stepCount++;
    }

    /* ---- refined energies (IntegrateRigidBodyStepKernel::getRefinedKineticEnergies
       and getPotentialEnergyRefinement).  The reference platform returns the plain
       values (ReferenceRigidBodyKernels.cpp:123-132); the CUDA platform's COMPMOD
       build computes these, restated here in FP64:
         - free atoms: posDot = s*D(-dt) + 5 D(dt) + 2 D'(dt), D(h) = (v + F/m h/2) h the
           displacement over h from the current state (freeAtomsDelta/freeAtomsDot,
           rigidbodyintegrator.cu:270-292, sequence CudaRigidBodyKernels.cpp:405-437),
           D' from the state after the step.  The reference uses s = -1; with
           RBREF_FIX_REFINED_FREE s = +1, the sign the body terms have;
         - bodies: rdot = 2.5 r(t) + (v - F/M dt/2) dt/2 before part 1 and
           rdot = 2.5 r(t+dt) + (v + F/M dt/2) dt - rdot after part 2; qdot =
           q_virt(-dt)/2 - 3 q(t), then qdot += 1.5 q(t+dt) + q_virt(+dt) and
           qdot -= q (qdot . q) (rigidbodyintegrator.cu:318-320, 380-383), q_virt(h) =
           q after the kick pi + T h and a free rotation over h (:238-243);
         - KE_t = [sum_free posDot.v m/2 + sum_bodies M rdot.v] / (6 dt),
           KE_r = sum_bodies qdot.pi / (6 dt) (rigidbodyintegrator.cu:433-446,
           CudaRigidBodyKernels.cpp:159-163);
         - U_ref = -dt^2/24 [sum_free |F|^2/m + sum_bodies |F|^2/M + tau.(I^-1 tau)],
           tau = B(q)^T T (rigidbodyintegrator.cu:449-469, CudaRigidBodyKernels.cpp:481-491).
       The accumulators are zeroed whenever the body system is rebuilt
       (uploadBodySystem, CudaRigidBodyKernels.cpp:293-335). */
    bool refined = false;
    std::vector<Vec3> posDot, posDelta;   /* per free atom */
    std::vector<Vec3> rdot;               /* per body */
    std::vector<Quat> qdot;

    void refineReset() {
        posDot.assign(kernelSystem.numFree, Vec3());
        posDelta.assign(kernelSystem.numFree, Vec3());
        rdot.assign(kernelSystem.numBodies, Vec3());
        qdot.assign(kernelSystem.numBodies, Quat());
    }
    Vec3 freeDelta(const FreeAtom& a, double h) const {
        return (ctx.V[a.index] + ctx.F[a.index] * ((0.5 * h) * a.invMass)) * h;
    }
    Quat virtualRotation(const Body& b, double h) const {
        Quat q = b.q, pi = b.pi + b.torque * h;
        rotateBody(q, pi, b.I, b.invI, b.dof, h, rotationMode, flags);
        return q;
    }
    void refineBegin() {
        const double halfDt = 0.5 * dt;
        const double s = (flags & RBREF_FIX_REFINED_FREE) ? 1.0 : -1.0;
        for (int k = 0; k < kernelSystem.numFree; k++) {
            const FreeAtom& a = kernelSystem.freeAtom[k];
            posDot[k] = freeDelta(a, -dt) * s;
            posDelta[k] = freeDelta(a, dt);
        }
        for (int k = 0; k < kernelSystem.numBodies; k++) {
            const Body& b = kernelSystem.body[k];
            Vec3 v = b.pcm * b.invMass;
            Vec3 dv = b.force * (b.invMass * halfDt);
            rdot[k] = b.rcm * 2.5 + (v - dv) * halfDt;
            qdot[k] = virtualRotation(b, -dt) * 0.5 - b.q * 3.0;
        }
    }
    void refineEnd() {
        const double halfDt = 0.5 * dt;
        for (int k = 0; k < kernelSystem.numFree; k++) {
            posDot[k] += posDelta[k] * 5.0;
            posDot[k] += freeDelta(kernelSystem.freeAtom[k], dt) * 2.0;
        }
        for (int k = 0; k < kernelSystem.numBodies; k++) {
            const Body& b = kernelSystem.body[k];
            Vec3 v = b.pcm * b.invMass;
            Vec3 dv = b.force * (b.invMass * halfDt);
            rdot[k] = b.rcm * 2.5 + (v + dv) * dt - rdot[k];
            qdot[k] += b.q * 1.5 + virtualRotation(b, dt);
            qdot[k] += b.q * (-qdot[k].dot(b.q));
        }
    }
    void refinedKineticEnergies(double* out2) const {
        double KEt = 0.0, KEr = 0.0;
        for (int k = 0; k < kernelSystem.numFree; k++) {
            const FreeAtom& a = kernelSystem.freeAtom[k];
            KEt += posDot[k].dot(ctx.V[a.index]) * (0.5 / a.invMass);
        }
        for (int k = 0; k < kernelSystem.numBodies; k++) {
            const Body& b = kernelSystem.body[k];
            KEt += rdot[k].dot(b.pcm * b.invMass) / b.invMass;
        }
        for (int k = 0; k < kernelSystem.numBodies; k++) KEr += qdot[k].dot(kernelSystem.body[k].pi);
        double factor = 1.0 / (6.0 * dt);
        out2[0] = KEt * factor;
        out2[1] = KEr * factor;
    }
    double potentialEnergyRefinement() const {
        double U = 0.0;
        for (const FreeAtom& a : kernelSystem.freeAtom) U += ctx.F[a.index].dot(ctx.F[a.index]) * a.invMass;
        for (const Body& b : kernelSystem.body) {
            Vec3 tau = b.q.Bt(b.torque);
            U += b.force.dot(b.force) * b.invMass + diagMul(tau, b.invI).dot(tau);
        }
        return -U * dt * dt / 24.0;
    }
};

namespace {
template <class F> int guarded(F&& f) {
    try { f(); return 0; }
    catch (const std::exception& e) { g_error = e.what(); return -1; }
}
void loadVec(std::vector<Vec3>& dst, const double* xyz, int n) {
    for (int i = 0; i < n; i++) dst[i] = Vec3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}
void storeVec(const std::vector<Vec3>& src, double* xyz, int n) {
    for (int i = 0; i < n; i++) for (int c = 0; c < 3; c++) xyz[3 * i + c] = src[i][c];
}
}  // namespace

extern "C" {

const char* rbref_last_error(void) { return g_error.c_str(); }

rbref_sim* rbref_create(int numAtoms, const double* masses, const int* bodyIndices, double dt,
                        int rotationMode, unsigned flags) {
    rbref_sim* s = new rbref_sim();
    int rc = guarded([&] {
        if (rotationMode < 0) throw std::runtime_error("Rotation mode cannot be negative");
        if (numAtoms <= 0) throw std::runtime_error("Number of body indices differs from that of atoms in Context");
        s->ctx.n = numAtoms;
        s->ctx.mass.assign(masses, masses + numAtoms);
        s->ctx.R.assign(numAtoms, Vec3());
        s->ctx.V.assign(numAtoms, Vec3());
        s->ctx.F.assign(numAtoms, Vec3());
        s->bodyIndices.assign(bodyIndices, bodyIndices + numAtoms);
        s->dt = dt;
        s->rotationMode = rotationMode;
        s->flags = flags;
        s->oldPos.assign(numAtoms, Vec3());
        s->invMass.resize(numAtoms);
        for (int i = 0; i < numAtoms; i++) s->invMass[i] = 1.0 / masses[i];
        s->integratorSystem.initialize(s->ctx, s->bodyIndices, rotationMode, flags);
        s->kernelSystem = s->integratorSystem;
    });
    if (rc != 0) { delete s; return nullptr; }
    return s;
}

void rbref_destroy(rbref_sim* s) { delete s; }

int rbref_set_tether(rbref_sim* s, const double* k, const double* r0) {
    return guarded([&] {
        s->ctx.tetherK.assign(k, k + s->ctx.n);
        s->ctx.tetherR0.assign(s->ctx.n, Vec3());
        loadVec(s->ctx.tetherR0, r0, s->ctx.n);
    });
}

int rbref_add_bond(rbref_sim* s, int a, int b, double length, double k) {
    return guarded([&] {
        if (a < 0 || b < 0 || a >= s->ctx.n || b >= s->ctx.n) throw std::runtime_error("bond atom out of range");
        s->ctx.bonds.push_back(Bond{a, b, length, k});
    });
}

int rbref_add_constraint(rbref_sim* s, int a, int b, double distance) {
    return guarded([&] {
        if (a < 0 || b < 0 || a >= s->ctx.n || b >= s->ctx.n || a == b)
            throw std::runtime_error("constraint atoms out of range");
        if (!(distance > 0.0)) throw std::runtime_error("constraint distance must be positive");
        /* RigidBodySystem::initialize, RigidBodySystem.cpp:107-112 */
        if (s->integratorSystem.bodyIndex[a] != 0 || s->integratorSystem.bodyIndex[b] != 0)
            throw std::runtime_error("Constraints involving rigid-body atoms are not allowed");
        if (s->ctx.mass[a] == 0.0 || s->ctx.mass[b] == 0.0)
            throw std::runtime_error("A constraint cannot involve a massless particle");
        s->constraints.list.push_back(Constraint{a, b, distance});
        s->constraints.build(s->ctx.n);
        int nc = (int)s->constraints.list.size();
        s->integratorSystem.numConstraints = s->kernelSystem.numConstraints = nc;
    });
}

int rbref_add_virtual_site(rbref_sim* s, int type, int site, const int* parents, const double* weights) {
    return guarded([&] {
        int n = s->ctx.n;
        int np = type == RBREF_SITE_TWO_AVERAGE ? 2 : 3;
        if (type < RBREF_SITE_TWO_AVERAGE || type > RBREF_SITE_LOCAL_COORDINATES)
            throw std::runtime_error("unknown site type");
        if (site < 0 || site >= n) throw std::runtime_error("site out of range");
        if (s->ctx.mass[site] != 0.0) throw std::runtime_error("a virtual site must be massless");
        VirtualSite v{type, site, {0, 0, 0}, {}};
        for (int k = 0; k < np; k++) {
            int p = parents[k];
            if (p < 0 || p >= n || p == site || s->ctx.mass[p] == 0.0)
                throw std::runtime_error("a site's parents must be massive particles other than the site");
            v.p[k] = p;
        }
        for (int k = 0; k < (type == RBREF_SITE_LOCAL_COORDINATES ? 12 : 3); k++) v.w[k] = weights[k];
        for (const VirtualSite& o : s->ctx.sites)
            if (o.site == site) throw std::runtime_error("particle is already a virtual site");
        s->ctx.sites.push_back(v);
        /* RigidBodySystem.cpp:65-68: virtual sites are not actual atoms */
        for (BodySystem* b : {&s->integratorSystem, &s->kernelSystem}) {
            b->numActualAtoms--;
            b->numBodyAtoms = b->numActualAtoms - b->numFree;
            b->atomIndex.resize(b->numActualAtoms);
        }
    });
}

int rbref_set_constraint_tolerance(rbref_sim* s, double tol) {
    return guarded([&] {
        if (!(tol > 0.0)) throw std::runtime_error("constraint tolerance must be positive");
        s->constraints.tol = tol;
    });
}

long rbref_constraint_failures(rbref_sim* s) { return s->constraints.failures; }

int rbref_set_positions(rbref_sim* s, const double* xyz) {
    return guarded([&] { loadVec(s->ctx.R, xyz, s->ctx.n); s->stateChanged(true); });
}

int rbref_set_velocities(rbref_sim* s, const double* xyz) {
    return guarded([&] { loadVec(s->ctx.V, xyz, s->ctx.n); s->stateChanged(false); });
}

int rbref_set_step_size(rbref_sim* s, double dt) { s->dt = dt; return 0; }

int rbref_step(rbref_sim* s, int steps) {
    return guarded([&] { for (int i = 0; i < steps; i++) s->execute(); });
}

int rbref_get_positions(rbref_sim* s, double* xyz) { storeVec(s->ctx.R, xyz, s->ctx.n); return 0; }
int rbref_get_velocities(rbref_sim* s, double* xyz) { storeVec(s->ctx.V, xyz, s->ctx.n); return 0; }
int rbref_get_forces(rbref_sim* s, double* xyz) { storeVec(s->ctx.F, xyz, s->ctx.n); return 0; }

int rbref_kinetic_energies(rbref_sim* s, double* out2) {
    s->kernelSystem.computeKineticEnergies(s->ctx.V);
    out2[0] = s->kernelSystem.transKE;
    out2[1] = s->kernelSystem.rotKE;
    return 0;
}

int rbref_set_refined(rbref_sim* s, int enable) {
    s->refined = enable != 0;
    s->refineReset();
    return 0;
}

int rbref_refined_kinetic_energies(rbref_sim* s, double* out2) {
    if (s->refined) s->refinedKineticEnergies(out2);
    else rbref_kinetic_energies(s, out2);
    return 0;
}

double rbref_potential_energy_refinement(rbref_sim* s) {
    return s->refined ? s->potentialEnergyRefinement() : 0.0;
}

double rbref_potential_energy(rbref_sim* s) { return s->ctx.pe; }
double rbref_time(rbref_sim* s) { return s->ctx.time; }

int rbref_counts(rbref_sim* s, int* out5) {
    const BodySystem& b = s->kernelSystem;
    out5[0] = b.numFree; out5[1] = b.numBodies; out5[2] = b.numActualAtoms;
    out5[3] = b.numBodyAtoms; out5[4] = b.numDOF;
    return 0;
}

int rbref_atom_index(rbref_sim* s, int* out) {
    std::copy(s->kernelSystem.atomIndex.begin(), s->kernelSystem.atomIndex.end(), out);
    return 0;
}

int rbref_body_index(rbref_sim* s, int* out) {
    std::copy(s->kernelSystem.bodyIndex.begin(), s->kernelSystem.bodyIndex.end(), out);
    return 0;
}

int rbref_body_state(rbref_sim* s, int k, double* o) {
    if (k < 0 || k >= s->kernelSystem.numBodies) { g_error = "body index out of range"; return -1; }
    const Body& b = s->kernelSystem.body[k];
    int p = 0;
    o[p++] = b.N; o[p++] = b.dof; o[p++] = b.mass;
    for (int c = 0; c < 3; c++) o[p++] = b.I[c];
    for (int c = 0; c < 3; c++) o[p++] = b.invI[c];
    for (int c = 0; c < 3; c++) o[p++] = b.rcm[c];
    for (int c = 0; c < 3; c++) o[p++] = b.pcm[c];
    for (int c = 0; c < 4; c++) o[p++] = b.q[c];
    for (int c = 0; c < 4; c++) o[p++] = b.pi[c];
    for (int c = 0; c < 3; c++) o[p++] = b.force[c];
    for (int c = 0; c < 4; c++) o[p++] = b.torque[c];
    o[p++] = b.twoKt; o[p++] = b.twoKr; o[p++] = b.loc;
    return 0;
}

int rbref_body_fixed_positions(rbref_sim* s, double* xyz) {
    storeVec(s->kernelSystem.d, xyz, s->kernelSystem.numBodyAtoms);
    return 0;
}

double rbref_carlson_rf(double x, double y, double z) { return carlsonRF(x, y, z); }
double rbref_carlson_rj(double x, double y, double z, double p) { return carlsonRJ(x, y, z, p); }
double rbref_carlson_rc(double x, double y) { return carlsonRC(x, y); }
double rbref_omega(double x, double n, double m) { return omegaFn(x, n, m); }
void rbref_jacobi(double u, double m, double* sn, double* cn, double* dn) { jacobiFn(u, m, *sn, *cn, *dn); }

void rbref_eigen(const double* A9, double* w3, double* V9) {
    Mat3 A(Vec3(A9[0], A9[1], A9[2]), Vec3(A9[3], A9[4], A9[5]), Vec3(A9[6], A9[7], A9[8]));
    Vec3 w = eigenvalues(A);
    Mat3 V = eigenvectors(A, w);
    for (int i = 0; i < 3; i++) {
        w3[i] = w[i];
        for (int j = 0; j < 3; j++) V9[3 * i + j] = V[i][j];
    }
}

void rbref_quat_from_matrix(const double* A9, double* q4) {
    Mat3 A(Vec3(A9[0], A9[1], A9[2]), Vec3(A9[3], A9[4], A9[5]), Vec3(A9[6], A9[7], A9[8]));
    Quat q = quatFromMatrix(A);
    for (int i = 0; i < 4; i++) q4[i] = q[i];
}

void rbref_quat_rotate(const double* q4, const double* x3, int transpose, double* y3) {
    Quat q(q4[0], q4[1], q4[2], q4[3]);
    Vec3 x(x3[0], x3[1], x3[2]);
    Vec3 y = transpose ? q.At(x) : q.A(x);
    for (int i = 0; i < 3; i++) y3[i] = y[i];
}

void rbref_clean_body_indices(int n, const int* in, int* out) {
    std::vector<int> r = cleanBodyIndices(std::vector<int>(in, in + n));
    std::copy(r.begin(), r.end(), out);
}

void rbref_rotate(double* q4, double* pi4, const double* I3, const double* invI3, int dof, double dt,
                  int mode, unsigned flags) {
    Quat q(q4[0], q4[1], q4[2], q4[3]), pi(pi4[0], pi4[1], pi4[2], pi4[3]);
    rotateBody(q, pi, Vec3(I3[0], I3[1], I3[2]), Vec3(invI3[0], invI3[1], invI3[2]), dof, dt, mode, flags);
    for (int i = 0; i < 4; i++) { q4[i] = q[i]; pi4[i] = pi[i]; }
}

}  /* extern "C" */
