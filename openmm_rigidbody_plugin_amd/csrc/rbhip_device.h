// Device math for the rigid-body step: quaternion algebra, free-rotor
// rotations (NO-SQUISH and exact), FP64 elliptic integrals, 3x3 symmetric
// eigen-decomposition.  gfx950, wave64.
//
// Quaternion convention (reference openmmapi/src/MatVec.cpp:495-549):
// q = (q0, q1, q2, q3); B(q) has columns q (x) e_k, C(q) columns e_k (x) q;
// A(q) = B(q)^T C(q) maps space -> body, A(q)^T = C(q)^T B(q) body -> space.
// The reference's quaternion momentum is pi = 2 B(q) L with L the body-frame
// angular momentum; the device state carries L, and a torque kick
// pi += C(q) tau dt is L += A(q) tau dt / 2.
#pragma once

#include <hip/hip_runtime.h>

namespace rbhip {

template <class T> struct V3 { T x, y, z; };
template <class T> struct Q4 { T a, b, c, d; };  // (q0, q1, q2, q3)

template <class T> __device__ __forceinline__ V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> __device__ __forceinline__ Q4<T> q4(T a, T b, T c, T d) { return Q4<T>{a, b, c, d}; }

template <class T> __device__ __forceinline__ V3<T> operator+(V3<T> u, V3<T> v) { return {u.x + v.x, u.y + v.y, u.z + v.z}; }
template <class T> __device__ __forceinline__ V3<T> operator-(V3<T> u, V3<T> v) { return {u.x - v.x, u.y - v.y, u.z - v.z}; }
template <class T> __device__ __forceinline__ V3<T> operator*(V3<T> u, T s) { return {u.x * s, u.y * s, u.z * s}; }
template <class T> __device__ __forceinline__ T dot(V3<T> u, V3<T> v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
template <class T> __device__ __forceinline__ V3<T> cross(V3<T> u, V3<T> v) {
    return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
template <class T> __device__ __forceinline__ V3<T> hadamard(V3<T> u, V3<T> v) { return {u.x * v.x, u.y * v.y, u.z * v.z}; }
template <class T> __device__ __forceinline__ Q4<T> operator+(Q4<T> u, Q4<T> v) { return {u.a + v.a, u.b + v.b, u.c + v.c, u.d + v.d}; }
template <class T> __device__ __forceinline__ Q4<T> operator*(Q4<T> u, T s) { return {u.a * s, u.b * s, u.c * s, u.d * s}; }
template <class T> __device__ __forceinline__ T dot(Q4<T> u, Q4<T> v) { return u.a * v.a + u.b * v.b + u.c * v.c + u.d * v.d; }

template <class To, class From> __device__ __forceinline__ V3<To> cvt(V3<From> v) { return {(To)v.x, (To)v.y, (To)v.z}; }
template <class To, class From> __device__ __forceinline__ Q4<To> cvt(Q4<From> v) { return {(To)v.a, (To)v.b, (To)v.c, (To)v.d}; }

// B(q) x
template <class T> __device__ __forceinline__ Q4<T> Bmul(Q4<T> q, V3<T> x) {
    return {-q.b * x.x - q.c * x.y - q.d * x.z,
             q.a * x.x - q.d * x.y + q.c * x.z,
             q.d * x.x + q.a * x.y - q.b * x.z,
            -q.c * x.x + q.b * x.y + q.a * x.z};
}
// C(q) x
template <class T> __device__ __forceinline__ Q4<T> Cmul(Q4<T> q, V3<T> x) {
    return {-q.b * x.x - q.c * x.y - q.d * x.z,
             q.a * x.x + q.d * x.y - q.c * x.z,
            -q.d * x.x + q.a * x.y + q.b * x.z,
             q.c * x.x - q.b * x.y + q.a * x.z};
}
// B(q)^T y
template <class T> __device__ __forceinline__ V3<T> Btmul(Q4<T> q, Q4<T> y) {
    return {-q.b * y.a + q.a * y.b + q.d * y.c - q.c * y.d,
            -q.c * y.a - q.d * y.b + q.a * y.c + q.b * y.d,
            -q.d * y.a + q.c * y.b - q.b * y.c + q.a * y.d};
}
// C(q)^T y
template <class T> __device__ __forceinline__ V3<T> Ctmul(Q4<T> q, Q4<T> y) {
    return {-q.b * y.a + q.a * y.b - q.d * y.c + q.c * y.d,
            -q.c * y.a + q.d * y.b + q.a * y.c - q.b * y.d,
            -q.d * y.a - q.c * y.b + q.b * y.c + q.a * y.d};
}
// P_k q = q (x) e_k
template <int K, class T> __device__ __forceinline__ Q4<T> perm(Q4<T> q) {
    if (K == 1) return {-q.b, q.a, q.d, -q.c};
    if (K == 2) return {-q.c, -q.d, q.a, q.b};
    return {-q.d, q.c, -q.b, q.a};
}
template <class T> __device__ __forceinline__ V3<T> Amul(Q4<T> q, V3<T> x) { return Btmul(q, Cmul(q, x)); }
template <class T> __device__ __forceinline__ V3<T> Atmul(Q4<T> q, V3<T> x) { return Ctmul(q, Bmul(q, x)); }

// The matrix of A(q)^T (body -> space), built once per body: rows r0..r2.
template <class T> struct Rot { V3<T> r0, r1, r2; };
template <class T> __device__ __forceinline__ Rot<T> rotT(Q4<T> q) {
    T aa = q.a * q.a, bb = q.b * q.b, cc = q.c * q.c, dd = q.d * q.d;
    T ab = q.a * q.b, ac = q.a * q.c, ad = q.a * q.d;
    T bc = q.b * q.c, bd = q.b * q.d, cd = q.c * q.d;
    Rot<T> R;
    R.r0 = {aa + bb - cc - dd, (T)2 * (bc - ad), (T)2 * (ac + bd)};
    R.r1 = {(T)2 * (ad + bc), aa - bb + cc - dd, (T)2 * (cd - ab)};
    R.r2 = {(T)2 * (bd - ac), (T)2 * (ab + cd), aa - bb - cc + dd};
    return R;
}
template <class T> __device__ __forceinline__ V3<T> apply(const Rot<T>& R, V3<T> x) {
    return {dot(R.r0, x), dot(R.r1, x), dot(R.r2, x)};
}

// ---------------------------------------------------------------------------
// sin(phi) and cos(phi) - 1.  The rotation below updates q <- q + q (cos-1)
// + P q sin, so the scale factor (1 + (cos-1))^2 + sin^2 stays 1 to the
// accuracy of (cos-1) itself rather than to the rounding of cos near 1.
// fp32: Taylor polynomials for |phi| <= 1/4 (truncation < 5e-11 relative);
// fp64: Taylor polynomials for |phi| <= 1/16; otherwise half-angle
// identities on the library sincos.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sin_cosm1(float phi, float& s, float& cm1) {
    if (fabsf(phi) <= 0.25f) {
        float x2 = phi * phi;
        s = phi * fmaf(x2, fmaf(x2, fmaf(x2, -1.0f / 5040.0f, 1.0f / 120.0f), -1.0f / 6.0f), 1.0f);
        cm1 = x2 * fmaf(x2, fmaf(x2, fmaf(x2, 1.0f / 40320.0f, -1.0f / 720.0f), 1.0f / 24.0f), -0.5f);
    } else {
        float sh, ch;
        sincosf(0.5f * phi, &sh, &ch);
        s = 2.0f * sh * ch;
        cm1 = -2.0f * sh * sh;
    }
}
__device__ __forceinline__ void sin_cosm1(double phi, double& s, double& cm1) {
    if (fabs(phi) <= 0.0625) {
        // Taylor polynomials through phi^9 / phi^10: truncation < 2e-20 relative
        double x2 = phi * phi;
        s = phi * fma(x2, fma(x2, fma(x2, fma(x2, 1.0 / 362880.0, -1.0 / 5040.0), 1.0 / 120.0), -1.0 / 6.0), 1.0);
        cm1 = x2 * fma(x2, fma(x2, fma(x2, fma(x2, -1.0 / 3628800.0, 1.0 / 40320.0), -1.0 / 720.0), 1.0 / 24.0), -0.5);
    } else {
        double sh, ch;
        sincos(0.5 * phi, &sh, &ch);
        s = 2.0 * sh * ch;
        cm1 = -2.0 * sh * sh;
    }
}

// sin(phi) and cos(phi) - 1 on |phi| <= PHI_FAST without a branch:
// fp32 Taylor through phi^5 / phi^6 (truncation < 3e-11 relative at 0.1),
// fp64 through phi^9 / phi^10 (< 2e-20 relative at 1/16).
__device__ __forceinline__ void sin_cosm1_small(float phi, float& s, float& cm1) {
    float x2 = phi * phi;
    s = phi * fmaf(x2, fmaf(x2, 1.0f / 120.0f, -1.0f / 6.0f), 1.0f);
    cm1 = x2 * fmaf(x2, fmaf(x2, -1.0f / 720.0f, 1.0f / 24.0f), -0.5f);
}
__device__ __forceinline__ void sin_cosm1_small(double phi, double& s, double& cm1) {
    double x2 = phi * phi;
    s = phi * fma(x2, fma(x2, fma(x2, fma(x2, 1.0 / 362880.0, -1.0 / 5040.0), 1.0 / 120.0), -1.0 / 6.0), 1.0);
    cm1 = x2 * fma(x2, fma(x2, fma(x2, fma(x2, -1.0 / 3628800.0, 1.0 / 40320.0), -1.0 / 720.0), 1.0 / 24.0), -0.5);
}
// fp32 on |phi| <= 0.03: through phi^3 / phi^4 (truncation < 7e-9 relative,
// below fp32 rounding)
__device__ __forceinline__ void sin_cosm1_tiny(float phi, float& s, float& cm1) {
    float x2 = phi * phi;
    s = fmaf(x2 * phi, -1.0f / 6.0f, phi);
    cm1 = x2 * fmaf(x2, 1.0f / 24.0f, -0.5f);
}
__device__ __forceinline__ void sin_cosm1_tiny(double phi, double& s, double& cm1) { sin_cosm1_small(phi, s, cm1); }
// sin and cos of an FP64 angle that is usually small (one step's turn): the
// Taylor polynomials above for |x| <= 1/16 (truncation < 2e-20 relative), the
// library's sincos otherwise
__device__ __forceinline__ void sincos_turn(double x, double& s, double& c) {
    if (fabs(x) <= 0.0625) {
        double cm1;
        sin_cosm1_small(x, s, cm1);
        c = 1.0 + cm1;
    } else {
        sincos(x, &s, &c);
    }
}
// 1 / sqrt(n2) for the squared norm of a quaternion being renormalised:
// 1 - e/2 + 3e^2/8 for |e| = |n2 - 1| < 1e-5 (truncation < 4e-16), the square
// root otherwise
__device__ __forceinline__ double inv_norm(double n2) {
    const double e = n2 - 1.0;
    return fabs(e) < 1e-5 ? 1.0 + e * (-0.5 + 0.375 * e) : 1.0 / sqrt(n2);
}
template <class T> struct SmallAngle;
template <> struct SmallAngle<float> { static constexpr float value = 0.1f; };
template <> struct SmallAngle<double> { static constexpr double value = 0.0625; };
// (FP64 has no shorter tier)
template <class T> struct TinyAngle;
template <> struct TinyAngle<float> { static constexpr float value = 0.03f; };
template <> struct TinyAngle<double> { static constexpr double value = -1.0; };

// the sin / cos - 1 evaluation a rotation uses: TIER 0 general, 1 the
// small-angle polynomials, 2 the tiny-angle ones
template <int TIER, class T> __device__ __forceinline__ void sin_cosm1_tier(T phi, T& s, T& cm1) {
    if (TIER == 2) sin_cosm1_tiny(phi, s, cm1);
    else if (TIER == 1) sin_cosm1_small(phi, s, cm1);
    else sin_cosm1(phi, s, cm1);
}

// One uniaxial rotation of NO-SQUISH in body-frame variables.  For
// pi = 2 B(q) L and |q| = 1 the reference's angle (pi . P_K q) tau / (4 I_K)
// (RigidBody.cpp:189-214) equals L_K tau / (2 I_K), so the angles depend on L
// only: L turns about body axis K by -2 phi, and the orientation update
// q <- q (x) (cos phi, sin phi e_K) is accumulated into a composite rotation
// r, applied to q once per step.
//
// Within a step r stays near the identity and L near its start value L0, so
// both are carried as offsets, r = 1 + rho and L = L0 + lam: their rounding
// errors scale with |rho| ~ w dt / 2 and |lam| ~ w dt |L| instead of with 1
// and |L|.  In fp32 this keeps the sequential updates of a 10-substep step
// from random-walking the orientation (~1e-6 rad per 10 steps when q itself is
// updated in fp32) at the cost of fp32 arithmetic.
// (coef and L0coef = L0_K coef: the angle is phi = L0coef + lam_K coef)
template <int K, int TIER, class T>
__device__ __forceinline__ void body_axis_turn(V3<T> L0, V3<T>& lam, Q4<T>& rho, T coef, T L0coef) {
    T lamK = K == 1 ? lam.x : (K == 2 ? lam.y : lam.z);
    T s, cm1;
    sin_cosm1_tier<TIER>(fma(lamK, coef, L0coef), s, cm1);
    // rho <- rho + rho cm1 + P_K(rho) s + (cm1, s e_K)   [r (x) (1 + cm1, s e_K)]
    // The e_K component of P_K(rho) is rho.a, so (cm1, s e_K) joins the terms
    // in rho.a as (1 + rho.a) cm1 and (1 + rho.a) s.  Rounding 1 + rho.a costs
    // at most 2^-24 |s|, below the rounding of rho itself over the step.
    const T a1 = rho.a + (T)1;
    Q4<T> Pr = perm<K>(rho);
    if (K == 1) Pr.b = a1;
    if (K == 2) Pr.c = a1;
    if (K == 3) Pr.d = a1;
    rho = {fma(a1, cm1, fma(Pr.a, s, rho.a)), fma(rho.b, cm1, fma(Pr.b, s, rho.b)),
           fma(rho.c, cm1, fma(Pr.c, s, rho.c)), fma(rho.d, cm1, fma(Pr.d, s, rho.d))};
    // L turns by -2 phi about axis K: sin 2phi = 2s (1 + cm1), cos 2phi - 1 = -2 s^2
    T s1 = s + s;
    T s2 = fma(s1, cm1, s1);
    T c2m1 = -(s1 * s);
    T& la = K == 1 ? lam.y : (K == 2 ? lam.z : lam.x);
    T& lb = K == 1 ? lam.z : (K == 2 ? lam.x : lam.y);
    T La = (K == 1 ? L0.y : (K == 2 ? L0.z : L0.x)) + la;
    T Lb = (K == 1 ? L0.z : (K == 2 ? L0.x : L0.y)) + lb;
    la = fma(La, c2m1, fma(Lb, s2, la));
    lb = fma(Lb, c2m1, fma(-La, s2, lb));
}

// The rotation sequence of NO-SQUISH(n) (reference RigidBody.cpp:220-231):
// n substeps of [R3(h/2)] R2(h/2) R1(h) R2(h/2) [R3(h/2)].  Rotations about one
// axis leave that component of L unchanged, so the back-to-back R3(h/2) R3(h/2)
// of neighbouring substeps are one R3(h) (R2 likewise for dof-5 bodies): 4n+1
// rotations instead of 5n (2n+1 instead of 3n).
template <int TIER, class T>
__device__ __forceinline__ void no_squish_turns(V3<T> L0, V3<T>& lam, Q4<T>& rho, V3<T> invI, T h, int n) {
    T c1 = (T)0.5 * h * invI.x;
    T c2h = (T)0.25 * h * invI.y, c2 = (T)0.5 * h * invI.y;
    T c3h = (T)0.25 * h * invI.z, c3 = (T)0.5 * h * invI.z;
    const T p1 = L0.x * c1, p2h = L0.y * c2h;
    if (invI.z != (T)0) {
        const T p3h = L0.z * c3h, p3 = L0.z * c3;
        body_axis_turn<3, TIER>(L0, lam, rho, c3h, p3h);
        for (int i = 0; i < n; i++) {
            body_axis_turn<2, TIER>(L0, lam, rho, c2h, p2h);
            body_axis_turn<1, TIER>(L0, lam, rho, c1, p1);
            body_axis_turn<2, TIER>(L0, lam, rho, c2h, p2h);
            const bool last = i + 1 == n;
            body_axis_turn<3, TIER>(L0, lam, rho, last ? c3h : c3, last ? p3h : p3);
        }
    } else {
        const T p2 = L0.y * c2;
        body_axis_turn<2, TIER>(L0, lam, rho, c2h, p2h);
        for (int i = 0; i < n; i++) {
            body_axis_turn<1, TIER>(L0, lam, rho, c1, p1);
            const bool last = i + 1 == n;
            body_axis_turn<2, TIER>(L0, lam, rho, last ? c2h : c2, last ? p2h : p2);
        }
    }
}

// NO-SQUISH rotation of (q, L) over dt with n substeps; dof-5 linear bodies
// are marked by invI.z == 0.  On return q = q (x) r renormalised in FP64 and
// L = L0 + lam formed in FP64.  |L| is conserved by every turn, so
// |phi| <= |L| max(coef) bounds all angles of the step and selects the
// branch-free small- or tiny-angle path for the whole body.
template <class T>
__device__ __forceinline__ void no_squish(Q4<T>& q, V3<T>& L, V3<T> invI, double dt, int n) {
    V3<T> L0 = L;
    V3<T> lam = {(T)0, (T)0, (T)0};
    Q4<T> rho = {(T)0, (T)0, (T)0, (T)0};
    T h = (T)(dt / (double)n);
    // (a bound, so fp32 is plenty: the 0.1% margin covers its rounding)
    float cmax = 0.5f * (float)h * fmaxf((float)invI.x, fmaxf((float)invI.y, (float)invI.z));
    V3<float> Lf = cvt<float>(L0);
    double phimax = 1.001f * sqrtf(dot(Lf, Lf)) * cmax;
    V3<double> Ld = cvt<double>(L0);
    if (phimax <= (double)TinyAngle<T>::value)
        no_squish_turns<2>(L0, lam, rho, invI, h, n);
    else if (phimax <= (double)SmallAngle<T>::value)
        no_squish_turns<1>(L0, lam, rho, invI, h, n);
    else
        no_squish_turns<0>(L0, lam, rho, invI, h, n);
    // q (x) r, renormalised in FP64
    Q4<double> qd = cvt<double>(q);
    Q4<double> r = {1.0 + (double)rho.a, (double)rho.b, (double)rho.c, (double)rho.d};
    Q4<double> qn = qd * r.a + perm<1>(qd) * r.b + perm<2>(qd) * r.c + perm<3>(qd) * r.d;
    double e = dot(qn, qn) - 1.0;
    double sc = fabs(e) < 1e-4 ? 1.0 + e * (-0.5 + 0.375 * e) : 1.0 / sqrt(1.0 + e);
    q = cvt<T>(qn * sc);
    L = cvt<T>(Ld + cvt<double>(lam));
}

// ---------------------------------------------------------------------------
// FP64 elliptic integrals by Carlson's duplication algorithms (Numer.
// Algorithms 10 (1995) 13-26, Algorithms 1 and 4, tolerance r = 1e-16), and
// Jacobi functions by Bulirsch's algorithm (Numer. Math. 7 (1965) 78-90):
// the arithmetic-geometric mean of (1, sqrt(1 - m)), which also gives
// K(m) = pi / (2 AGM) (DLMF 19.8.5), then one sin/cos and a backward
// recurrence with one division per level.
// ---------------------------------------------------------------------------
// R_C(1, 1 + e) in closed form (DLMF 19.2.17-19.2.18): atan(sqrt(e))/sqrt(e)
// for e > 0, atanh(sqrt(-e))/sqrt(-e) for -1 < e < 0 (1 - sqrt(-e) written
// as (1 + e)/(1 + sqrt(-e)) against cancellation).  Not inlined: the caller
// takes this branch rarely and keeps its own registers free of it.
__device__ __noinline__ double carlson_rc_one_closed(double e) {
    if (e > 0.0) {
        double t = sqrt(e);
        return atan(t) / t;
    }
    double t = sqrt(-e);
    return 0.5 * log1p(2.0 * t * (1.0 + t) / (1.0 + e)) / t;
}

// R_C(1, 1 + e): Taylor series 1 - e/3 + e^2/5 - e^3/7 + e^4/9 for |e| < 1e-3
// (truncation below 1e-16), the closed form otherwise.  In the duplication
// loop of R_J, e shrinks by about 64 per step, so after the first steps the
// series applies.
__device__ __forceinline__ double carlson_rc_one(double e) {
    if (fabs(e) < 1e-3) return 1.0 + e * (-1.0 / 3.0 + e * (1.0 / 5.0 + e * (-1.0 / 7.0 + e * (1.0 / 9.0))));
    return carlson_rc_one_closed(e);
}

__device__ inline double carlson_rj(double x, double y, double z, double p) {
    double A0 = (x + y + z + 2.0 * p) * 0.2;
    double delta = (p - x) * (p - y) * (p - z);
    double Q = 584.803547642573 *   // (2.5e-17)^(-1/6)
               fmax(fmax(fabs(A0 - x), fabs(A0 - y)), fmax(fabs(A0 - z), fabs(A0 - p)));
    double xm = x, ym = y, zm = z, pm = p, Am = A0, scale = 1.0, scale3 = 1.0, sum = 0.0;
    for (int m = 0; m < 60 && scale * Q >= fabs(Am); m++) {
        double sx = sqrt(xm), sy = sqrt(ym), sz = sqrt(zm), sp = sqrt(pm);
        double lam = sx * sy + sx * sz + sy * sz;
        double dm = (sp + sx) * (sp + sy) * (sp + sz);
        double e = scale3 * delta / (dm * dm);
        sum += scale * carlson_rc_one(e) / dm;
        xm = 0.25 * (xm + lam);
        ym = 0.25 * (ym + lam);
        zm = 0.25 * (zm + lam);
        pm = 0.25 * (pm + lam);
        Am = 0.25 * (Am + lam);
        scale *= 0.25;
        scale3 *= 1.0 / 64.0;
    }
    double X = (A0 - x) * scale / Am, Y = (A0 - y) * scale / Am, Z = (A0 - z) * scale / Am;
    double P = -0.5 * (X + Y + Z);
    double E2 = X * Y + X * Z + Y * Z - 3.0 * P * P;
    double E3 = X * Y * Z + 2.0 * E2 * P + 4.0 * P * P * P;
    double E4 = (2.0 * X * Y * Z + E2 * P + 3.0 * P * P * P) * P;
    double E5 = X * Y * Z * P * P;
    double poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 -
                  9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    return scale * poly / (Am * sqrt(Am)) + 6.0 * sum;
}

// R_F(x, y, z) and R_J(x, y, z, p) from one duplication sequence: both
// algorithms duplicate (x, y, z) with the same lambda (R_J adds p), so one
// loop serves both; it runs until both meet their tolerances
__device__ inline void carlson_rf_rj(double x, double y, double z, double p, double& rf, double& rj) {
    double AF0 = (x + y + z) * (1.0 / 3.0);
    double AJ0 = (x + y + z + 2.0 * p) * 0.2;
    double delta = (p - x) * (p - y) * (p - z);
    double QF = 386.49729393434575 * fmax(fabs(AF0 - x), fmax(fabs(AF0 - y), fabs(AF0 - z)));  // (3e-16)^(-1/6)
    double QJ = 584.803547642573 *   // (2.5e-17)^(-1/6)
                fmax(fmax(fabs(AJ0 - x), fabs(AJ0 - y)), fmax(fabs(AJ0 - z), fabs(AJ0 - p)));
    double xm = x, ym = y, zm = z, pm = p, AF = AF0, AJ = AJ0, scale = 1.0, scale3 = 1.0, sum = 0.0;
    for (int m = 0; m < 60 && (scale * QJ >= fabs(AJ) || scale * QF >= fabs(AF)); m++) {
        double sx = sqrt(xm), sy = sqrt(ym), sz = sqrt(zm), sp = sqrt(pm);
        double lam = sx * sy + sx * sz + sy * sz;
        double dm = (sp + sx) * (sp + sy) * (sp + sz);
        double e = scale3 * delta / (dm * dm);
        sum += scale * carlson_rc_one(e) / dm;
        xm = 0.25 * (xm + lam);
        ym = 0.25 * (ym + lam);
        zm = 0.25 * (zm + lam);
        pm = 0.25 * (pm + lam);
        AF = 0.25 * (AF + lam);
        AJ = 0.25 * (AJ + lam);
        scale *= 0.25;
        scale3 *= 1.0 / 64.0;
    }
    {
        double X = (AF0 - x) * scale / AF, Y = (AF0 - y) * scale / AF, Z = -(X + Y);
        double E2 = X * Y - Z * Z, E3 = X * Y * Z;
        rf = (1.0 - E2 / 10.0 + E3 / 14.0 + E2 * E2 / 24.0 - 3.0 * E2 * E3 / 44.0) / sqrt(AF);
    }
    double X = (AJ0 - x) * scale / AJ, Y = (AJ0 - y) * scale / AJ, Z = (AJ0 - z) * scale / AJ;
    double P = -0.5 * (X + Y + Z);
    double E2 = X * Y + X * Z + Y * Z - 3.0 * P * P;
    double E3 = X * Y * Z + 2.0 * E2 * P + 4.0 * P * P * P;
    double E4 = (2.0 * X * Y * Z + E2 * P + 3.0 * P * P * P) * P;
    double E5 = X * Y * Z * P * P;
    double poly = 1.0 - 3.0 * E2 / 14.0 + E3 / 6.0 + 9.0 * E2 * E2 / 88.0 - 3.0 * E4 / 22.0 -
                  9.0 * E2 * E3 / 52.0 + 3.0 * E5 / 26.0;
    rj = scale * poly / (AJ * sqrt(AJ)) + 6.0 * sum;
}

// Omega(x, n, m) = -(n x^3 / 3) R_J(1 - x^2, 1 - m x^2, 1, 1 + n x^2)
// (the function used by the exact rotation, reference ellipticFunctions.h:251-254)
__device__ inline double omega_fn(double x, double n, double m) {
    double x2 = x * x;
    return (-1.0 / 3.0) * n * x * x2 * carlson_rj(1.0 - x2, 1.0 - m * x2, 1.0, 1.0 + n * x2);
}

// The AGM levels of (1, sqrt(1 - m)), kept for the Jacobi functions; the
// fixed trip count keeps the arrays in registers.  The mean c is exact to
// about CA^2 once |a - b| <= CA a; 8 levels reach that for any 1 - m >= 2 eps
// (smaller 1 - m takes the m = 1 limit in jacobi_agm).
constexpr int kAgmLevels = 9;
struct AgmLevels {
    double a[kAgmLevels], b[kAgmLevels];
    double c;    // the arithmetic-geometric mean
    int last;    // index of the last level used
};
__device__ inline void agm_levels(double m, AgmLevels& g) {
    const double CA = 1.0e-8;
    double a = 1.0, bb = 1.0 - m, c = 1.0;
    int last = 0;
    bool alive = true;
#pragma unroll
    for (int i = 0; i < kAgmLevels; i++) {
        if (alive) {
            g.a[i] = a;
            bb = sqrt(bb);
            g.b[i] = bb;
            c = 0.5 * (a + bb);
            last = i;
            if (fabs(a - bb) <= CA * a) {
                alive = false;
            } else {
                bb *= a;
                a = c;
            }
        }
    }
    g.c = c;
    g.last = last;
}

// sn, cn, dn of (u | m) from the AGM levels of m (Bulirsch's recurrence)
__device__ inline void jacobi_agm(double u, double m, const AgmLevels& g, double& sn, double& cn, double& dn) {
    const double eps = 2.220446049250313e-16;
    if (1.0 - m < 2.0 * eps) { sn = tanh(u); cn = 1.0 / cosh(u); dn = cn; return; }
    double su, cu;
    sincos(u * g.c, &su, &cu);
    sn = su;
    cn = cu;
    dn = 1.0;
    if (su == 0.0) return;
    double a = cu / su, c = g.c * a;
#pragma unroll
    for (int i = kAgmLevels - 1; i >= 0; i--) {
        if (i <= g.last) {
            double b = g.a[i];
            a *= c;
            c *= dn;
            dn = (g.b[i] + a) / (b + a);
            a = c / b;
        }
    }
    a = 1.0 / sqrt(c * c + 1.0);
    sn = su >= 0.0 ? a : -a;
    cn = c * sn;
}

__device__ __forceinline__ int sign_of(double x) { return x >= 0.0 ? 1 : -1; }
__device__ __forceinline__ int stair(double x) { return x > 0.0 ? (int)ceil(x - 0.5) : (int)floor(x + 0.5); }

// free linear rotor (dof 5): body-frame L constant, rotation about L at
// |w| = |L| / I0 (closed form; replaces the exact solution, which divides by
// I0 - I1 = 0 for these bodies)
__device__ inline void linear_rotor_exact(Q4<double>& q, V3<double> L, V3<double> invI, double dt) {
    V3<double> w = hadamard(invI, L);
    double wn = sqrt(dot(w, w));
    if (wn == 0.0) return;
    double sh, ch;
    sincos_turn(0.5 * wn * dt, sh, ch);
    V3<double> axis = w * (1.0 / wn);
    q = q * ch + Bmul(q, axis) * sh;
    q = q * inv_norm(dot(q, q));
}

// exact free-rotor rotation of (q, L) over dt by the reference's closed form
// (RigidBody.cpp:241-308), FP64; L is the body-frame angular momentum (the
// reference's B(q)^T pi / 2).  The reference's two branches on l1 < l3
// (:286-300) differ only in their parameters, so they are one computation
// with selected parameters here (the lanes of a wave then do not run both);
// K(m) comes from the AGM that the Jacobi functions need anyway, and
// F(s0 | m) (:270) and Omega(s0) (:289, :297) share one duplication sequence.
__device__ inline void exact_rotation_closed(Q4<double>& q, V3<double>& Lb, V3<double> I, V3<double> invI,
                                                   double dt) {
    const double eps = 2.220446049250313e-16;
    const double PI = 3.14159265358979323846;
    V3<double> Iw = Lb;
    V3<double> w0 = hadamard(invI, Iw);
    double Lsq = Iw.y * Iw.y + Iw.z * Iw.z;
    if (Lsq < eps) {
        // rotation about body axis 1 only (RigidBody.cpp:246-247): q turns by
        // theta = L1 dt / I1 about e1, L by -theta about e1
        double theta = Iw.x * invI.x * dt;
        double sh, ch, s2, c2;
        sincos(0.5 * theta, &sh, &ch);
        sincos(theta, &s2, &c2);
        q = q * ch + perm<1>(q) * sh;
        Lb = {Iw.x, Iw.y * c2 + Iw.z * s2, -Iw.y * s2 + Iw.z * c2};
        return;
    }
    Lsq += Iw.x * Iw.x;
    double L = sqrt(Lsq);
    double twoKr = dot(Iw, w0);
    Q4<double> z0 = {Iw.z, Iw.y, L - Iw.x, 0.0};
    double r1 = Lsq - twoKr * I.z;
    double r3 = twoKr * I.x - Lsq;
    double l1 = r1 * invI.y / (I.y - I.z);
    double l3 = r3 * invI.y / (I.x - I.y);
    const bool low = l1 < l3;
    double lmin = fmin(l1, l3);
    double c13 = 1.0 / (I.x - I.z);
    V3<double> a = {sign_of(w0.x) * sqrt(r1 * invI.x * c13), sqrt(lmin), sign_of(w0.z) * sqrt(r3 * invI.z * c13)};
    double m = lmin / fmax(l1, l3);
    AgmLevels g;
    agm_levels(m, g);
    double K = (0.5 * PI) / g.c;
    double inv2K = 0.5 / K;
    double alpha = I.x * a.x / L;
    double eta = alpha * alpha;
    eta /= 1.0 - eta;
    double npar = low ? eta : m * eta;   // the characteristic of Omega in each branch
    double s0 = w0.y / a.y;
    const bool inside = fabs(s0) < 1.0;
    if (!inside) {
        a.y = fabs(w0.y);
        s0 = sign_of(s0);
    }
    double S0 = s0 * s0, rf0, rj0;
    carlson_rf_rj(1.0 - S0, 1.0 - m * S0, 1.0, 1.0 + npar * S0, rf0, rj0);
    double c0 = inside ? (low ? w0.x / a.x : w0.z / a.z) : 0.0;
    double u0 = inside ? s0 * rf0 : s0 * K;
    int i0 = inside ? stair(u0 * inv2K) : 0;
    double wp = -invI.y * a.x * a.z / (a.y * c13);
    double u = wp * dt + u0;
    int jump = stair(u * inv2K) - i0;
    double sn, cn, dn;
    jacobi_agm(u, m, g, sn, cn, dn);
    V3<double> Ia = hadamard(I, a);
    double C = sqrt(low ? m + eta : 1.0 + m * eta);
    double om0 = (-1.0 / 3.0) * npar * s0 * S0 * rj0;
    double deltaF = u - u0 + sign_of(cn) * omega_fn(sn, npar, m) - sign_of(c0) * om0 +
                    (alpha / C) * (atan(C * sn / (low ? dn : cn)) - atan(C * s0 * a.z / w0.z));
    if (jump != 0) deltaF += jump * (2.0 * omega_fn(1.0, npar, m) + (low ? 0.0 : (alpha / C) * PI));
    Iw = low ? hadamard(Ia, V3<double>{cn, sn, dn}) : hadamard(Ia, V3<double>{dn, sn, cn});
    deltaF *= 1.0 + eta;
    double theta = (Lsq * (u - u0) + r3 * deltaF) / (2.0 * L * I.x * wp);
    double st, ct;
    sincos(theta, &st, &ct);
    Q4<double> z = Q4<double>{Iw.z, Iw.y, L - Iw.x, 0.0} * ct + Q4<double>{-Iw.y, Iw.z, 0.0, L - Iw.x} * st;
    q = z * dot(z0, q) + Cmul(z, Ctmul(z0, q));
    double r = 1.0 / sqrt(dot(q, q));
    q = q * r;
    Lb = Iw;
}

// Taylor coefficients of the Jacobi functions in u, each a polynomial in m
// (generated exactly by gen_jacobi_series.py from sn' = cn dn, cn' = -sn dn,
// dn' = -m sn cn): odd powers 1..17 of sn, even powers 0..16 of cn and dn.
// The truncated series are exact to 2e-16 for |u| <= 0.2 and any 0 <= m < 1.
__device__ constexpr double kSnCoef[9][9] = {
    {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {-0.16666666666666666, -0.16666666666666666, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.008333333333333333, 0.11666666666666667, 0.008333333333333333, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {-0.0001984126984126984, -0.026785714285714284, -0.026785714285714284, -0.0001984126984126984, 0.0, 0.0, 0.0, 0.0, 0.0},
    {2.7557319223985893e-06, 0.0033840388007054674, 0.015095899470899472, 0.0033840388007054674, 2.7557319223985893e-06, 0.0, 0.0, 0.0, 0.0},
    {-2.505210838544172e-08, -0.0002773017877184544, -0.004154290925124258, -0.004154290925124258, -0.0002773017877184544, -2.505210838544172e-08, 0.0, 0.0, 0.0},
    {1.6059043836821613e-10, 1.6001552459885793e-05, 0.0007217497972706306, 0.0021166250159305715, 0.0007217497972706306, 1.6001552459885793e-05, 1.6059043836821613e-10, 0.0, 0.0},
    {-7.647163731819816e-13, -6.857999376187207e-07, -8.893244055694453e-05, -0.0006382989522663795, -0.0006382989522663795, -8.893244055694453e-05, -6.857999376187207e-07, -7.647163731819816e-13, 0.0},
    {2.8114572543455206e-15, 2.269199123287981e-08, 8.293700555544529e-06, 0.00013256618181982426, 0.0003082622922067598, 0.00013256618181982426, 8.293700555544529e-06, 2.269199123287981e-08, 2.8114572543455206e-15},
};
__device__ constexpr double kCnCoef[9][9] = {
    {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {-0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.041666666666666664, 0.16666666666666666, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {-0.001388888888888889, -0.06111111111111111, -0.022222222222222223, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {2.48015873015873e-05, 0.01011904761904762, 0.02261904761904762, 0.0015873015873015873, 0.0, 0.0, 0.0, 0.0, 0.0},
    {-2.755731922398589e-07, -0.0010163139329805997, -0.008478835978835978, -0.004356261022927689, -7.054673721340387e-05, 0.0, 0.0, 0.0, 0.0},
    {2.08767569878681e-09, 6.933588530810753e-05, 0.0018176139703917481, 0.0032120143231254344, 0.0005413927636149858, 2.1377799155576935e-06, 0.0, 0.0, 0.0},
    {-1.1470745597729725e-11, -3.428972923020542e-06, -0.00026319690010166203, -0.0012264866233120202, -0.0007457005076052695, -4.79590955781432e-05, -4.698417396830095e-08, 0.0, 0.0},
    {4.779477332387385e-14, 1.285878228669234e-07, 2.8037847110267745e-05, 0.00030290726934245453, 0.0004712798102348367, 0.00012125411331760538, 3.2045164320296596e-06, 7.830695661383492e-10, 0.0},
};
__device__ constexpr double kDnCoef[9][9] = {
    {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, -0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.16666666666666666, 0.041666666666666664, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, -0.022222222222222223, -0.06111111111111111, -0.001388888888888889, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0015873015873015873, 0.02261904761904762, 0.01011904761904762, 2.48015873015873e-05, 0.0, 0.0, 0.0, 0.0},
    {0.0, -7.054673721340387e-05, -0.004356261022927689, -0.008478835978835978, -0.0010163139329805997, -2.755731922398589e-07, 0.0, 0.0, 0.0},
    {0.0, 2.1377799155576935e-06, 0.0005413927636149858, 0.0032120143231254344, 0.0018176139703917481, 6.933588530810753e-05, 2.08767569878681e-09, 0.0, 0.0},
    {0.0, -4.698417396830095e-08, -4.79590955781432e-05, -0.0007457005076052695, -0.0012264866233120202, -0.00026319690010166203, -3.428972923020542e-06, -1.1470745597729725e-11, 0.0},
    {0.0, 7.830695661383492e-10, 3.2045164320296596e-06, 0.00012125411331760538, 0.0004712798102348367, 0.00030290726934245453, 2.8037847110267745e-05, 1.285878228669234e-07, 4.779477332387385e-14},
};

struct JacobiSeries { double s[9], c[9], d[9]; };   // the coefficients at one m

// The first TERMS coefficients at m.  Row k of each table is a polynomial of
// degree k in m (k - 1 for cn), so Horner starts at that degree: the higher
// steps would be fma(0, m, 0), which is +0 for every finite m, so the values
// are the same bit for bit, and the compiler cannot drop those steps itself
// (0 * m is not 0 for m = inf or NaN).
template <int TERMS>
__device__ inline void jacobi_series_at(double m, JacobiSeries& J) {
#pragma unroll
    for (int k = 0; k < TERMS; k++) {
        double ps = kSnCoef[k][k], pd = kDnCoef[k][k];
#pragma unroll
        for (int j = k - 1; j >= 0; j--) {
            ps = fma(ps, m, kSnCoef[k][j]);
            pd = fma(pd, m, kDnCoef[k][j]);
        }
        const int kc = k > 0 ? k - 1 : 0;
        double pc = kCnCoef[k][kc];
#pragma unroll
        for (int j = kc - 1; j >= 0; j--) pc = fma(pc, m, kCnCoef[k][j]);
        J.s[k] = ps;
        J.c[k] = pc;
        J.d[k] = pd;
    }
}

// the series truncated after TERMS terms: order 2 TERMS - 1 for sn
template <int TERMS>
__device__ inline void jacobi_series(double u, const JacobiSeries& J, double& sn, double& cn, double& dn) {
    double u2 = u * u;
    double ps = J.s[TERMS - 1], pc = J.c[TERMS - 1], pd = J.d[TERMS - 1];
#pragma unroll
    for (int k = TERMS - 2; k >= 0; k--) {
        ps = fma(ps, u2, J.s[k]);
        pc = fma(pc, u2, J.c[k]);
        pd = fma(pd, u2, J.d[k]);
    }
    sn = u * ps;
    cn = pc;
    dn = pd;
}

// exact free-rotor rotation of (q, L) over dt, FP64.  The reference's solution
// is L(t) = I a (cn, sn, dn)(u0 + wp t) (or (dn, sn, cn)) and
// q(t + dt) = q (x) z0^* (x) z(theta) (RigidBody.cpp:251, 302-305).  When the
// elliptic argument advances little in one step, |wp dt| <= 0.2 (at 300 K and
// dt <= 2 fs it stays below 0.09), the same solution is evaluated without
// elliptic integrals:
//  - the Jacobi functions at u0 + x follow from their values at u0, which are
//    the angular velocity's components over a (s0, c0, d0), and their values
//    at x (the series above) by the addition theorem (DLMF 22.8.1-3);
//  - theta satisfies 2 dtheta/dt = (|L| + L1) X - L1 / I1 with
//    X = (L2^2/I2 + L3^2/I3) / (L2^2 + L3^2): q (x) n(t)^* stays constant for
//    the unit frame quaternion n of z(theta), and Euler's equations give its
//    rate; five-point Gauss-Legendre quadrature integrates it over the step.
// Against a tight ODE integration of Euler's equations this is exact to
// ~1e-15, where the closed form loses up to ~1e-11 near the axis-1 pole and
// the separatrix.  Longer steps take the closed form.
// FP32_STATE: the caller keeps q and L in fp32 (relative rounding 6e-8), so
// the series stops at order 11 and the quadrature takes three points; the
// result stays within 2e-13 of the full evaluation for |wp dt| <= 0.09 and
// 1e-10 up to 0.2, far below the rounding of the state it is stored in.
template <bool FP32_STATE = false>
__device__ inline void exact_rotation(Q4<double>& q, V3<double>& Lb, V3<double> I, V3<double> invI, double dt) {
    constexpr int TERMS = FP32_STATE ? 6 : 9;
    const double eps = 2.220446049250313e-16;
    // the frame quaternion z(L) is singular for L along +e1, and so is the
    // rate of theta in the complex neighbourhood of a step that passes near
    // it; a body with L1 > 0 is turned by pi about e3 first (q -> P3 q,
    // L -> (-L1, -L2, L3): the inertia tensor is unchanged) and back after
    const bool flip = Lb.x > 0.0;
    V3<double> Iw = flip ? V3<double>{-Lb.x, -Lb.y, Lb.z} : Lb;
    V3<double> w0 = hadamard(invI, Iw);
    double Lsq = Iw.y * Iw.y + Iw.z * Iw.z;
    if (Lsq < eps) {
        exact_rotation_closed(q, Lb, I, invI, dt);
        return;
    }
    Lsq += Iw.x * Iw.x;
    double L = sqrt(Lsq);
    double twoKr = dot(Iw, w0);
    double r1 = Lsq - twoKr * I.z;
    double r3 = twoKr * I.x - Lsq;
    double l1 = r1 * invI.y / (I.y - I.z);
    double l3 = r3 * invI.y / (I.x - I.y);
    const bool low = l1 < l3;
    double lmin = fmin(l1, l3);
    double c13 = 1.0 / (I.x - I.z);
    V3<double> a = {sign_of(w0.x) * sqrt(r1 * invI.x * c13), sqrt(lmin), sign_of(w0.z) * sqrt(r3 * invI.z * c13)};
    double m = lmin / fmax(l1, l3);
    double s0 = w0.y / a.y;
    const bool inside = fabs(s0) < 1.0;
    if (!inside) {
        a.y = fabs(w0.y);
        s0 = sign_of(s0);
    }
    double wp = -invI.y * a.x * a.z / (a.y * c13);
    double delta = wp * dt;
    if (!(fabs(delta) <= 0.2) || !(m >= 0.0 && m < 1.0)) {
        exact_rotation_closed(q, Lb, I, invI, dt);
        return;
    }
    double c0 = inside ? (low ? w0.x / a.x : w0.z / a.z) : 0.0;
    double d0 = inside ? (low ? w0.z / a.z : w0.x / a.x) : sqrt(1.0 - m);
    V3<double> Ia = hadamard(I, a);
    JacobiSeries J;
    jacobi_series_at<TERMS>(m, J);
    const double ms0 = m * s0, c0d0 = c0 * d0;
    // body-frame L at u0 + x as numerators over the addition theorem's
    // common denominator: L = Ln / D
    auto body_Ln = [&](double x, double& D) {
        double sx, cx, dx;
        jacobi_series<TERMS>(x, J, sx, cx, dx);
        D = 1.0 - ms0 * s0 * sx * sx;
        double sn = s0 * cx * dx + sx * c0d0;
        double cn = c0 * cx - s0 * sx * d0 * dx;
        double dn = d0 * dx - ms0 * sx * c0 * cx;
        return low ? V3<double>{Ia.x * cn, Ia.y * sn, Ia.z * dn} : V3<double>{Ia.x * dn, Ia.y * sn, Ia.z * cn};
    };
    auto body_L = [&](double x) {
        double D;
        V3<double> Ln = body_Ln(x, D);
        return Ln * (1.0 / D);
    };
    // 2 dtheta/dt at u0 + x, (|L| + L1) X - L1 / I1: X is the same for Ln as
    // for L, so one reciprocal serves X and L1 = Ln1 / D
    auto rate = [&](double x) {
        double D;
        V3<double> Ln = body_Ln(x, D);
        double y2 = Ln.y * Ln.y, z2 = Ln.z * Ln.z, s2 = y2 + z2;
        double p = 1.0 / (D * s2);
        double X = (y2 * invI.y + z2 * invI.z) * D * p;
        double L1 = Ln.x * s2 * p;
        return (L + L1) * X - L1 * invI.x;
    };
    // Gauss-Legendre nodes and weights on [-1, 1]: 5 points, or 3 for fp32 state
    const double h = 0.5 * delta;
    double sum;
    if constexpr (FP32_STATE) {
        const double gx1 = 0.7745966692414834;
        const double gw0 = 0.8888888888888888, gw1 = 0.5555555555555556;
        sum = gw0 * rate(h) + gw1 * (rate(h * (1.0 - gx1)) + rate(h * (1.0 + gx1)));
    } else {
        const double gx1 = 0.5384693101056831, gx2 = 0.9061798459386640;
        const double gw0 = 0.5688888888888889, gw1 = 0.4786286704993665, gw2 = 0.2369268850561891;
        sum = gw0 * rate(h) + gw1 * (rate(h * (1.0 - gx1)) + rate(h * (1.0 + gx1))) +
              gw2 * (rate(h * (1.0 - gx2)) + rate(h * (1.0 + gx2)));
    }
    double theta = 0.25 * dt * sum;
    V3<double> Iw1 = body_L(delta);
    Q4<double> z0 = {Iw.z, Iw.y, L - Iw.x, 0.0};
    double st, ct;
    sincos_turn(theta, st, ct);
    Q4<double> z = Q4<double>{Iw1.z, Iw1.y, L - Iw1.x, 0.0} * ct + Q4<double>{-Iw1.y, Iw1.z, 0.0, L - Iw1.x} * st;
    Q4<double> qf = flip ? perm<3>(q) : q;
    qf = z * dot(z0, qf) + Cmul(z, Ctmul(z0, qf));
    double r = inv_norm(dot(qf, qf));
    // P3 P3 q = -q: the turn back is -P3
    q = flip ? perm<3>(qf) * -r : qf * r;
    Lb = flip ? V3<double>{-Iw1.x, -Iw1.y, Iw1.z} : Iw1;
}

// ---------------------------------------------------------------------------
// Setup math in FP64 (body frame from atom positions), written to follow the
// reference's operation order; callers disable FMA contraction.
// Smith (1961) eigenvalues, descending (eigenDecomposition.cpp:72-100);
// Kopp (2008) cross-product eigenvectors (eigenDecomposition.cpp:104-176);
// Shepperd (1978) quaternion (MatVec.cpp:344-373).
// ---------------------------------------------------------------------------
struct M3 { V3<double> r[3]; };

__device__ inline double m3_det(const M3& A) {
    const V3<double>& a0 = A.r[0];
    const V3<double>& a1 = A.r[1];
    const V3<double>& a2 = A.r[2];
    return a0.x * (a1.y * a2.z - a2.y * a1.z) - a0.y * (a1.x * a2.z - a2.x * a1.z) +
           a0.z * (a1.x * a2.y - a2.x * a1.y);
}

__device__ inline double& m3_at(M3& A, int i, int j) {
    V3<double>& r = A.r[i];
    return j == 0 ? r.x : (j == 1 ? r.y : r.z);
}

__device__ inline V3<double> eigenvalues3(const M3& A) {
    const double eps = 2.220446049250313e-16;
    const double PI = 3.14159265358979323846;
    double p1 = A.r[0].y * A.r[0].y + A.r[0].z * A.r[0].z + A.r[1].z * A.r[1].z;
    if (p1 < eps) {
        double w0 = A.r[0].x, w1 = A.r[1].y, w2 = A.r[2].z, t;
        if (w0 < w1) { t = w0; w0 = w1; w1 = t; }
        if (w0 < w2) { t = w0; w0 = w2; w2 = t; }
        if (w1 < w2) { t = w1; w1 = w2; w2 = t; }
        return {w0, w1, w2};
    }
    double tr = A.r[0].x + A.r[1].y + A.r[2].z;
    double q = tr / 3.0;
    V3<double> d = {A.r[0].x - q, A.r[1].y - q, A.r[2].z - q};
    double p2 = dot(d, d) + 2.0 * p1;
    double p = sqrt(p2 / 6.0);
    M3 B = A;
    B.r[0].x -= q; B.r[1].y -= q; B.r[2].z -= q;
    double r = m3_det(B) * (3.0 / (p * p2));
    double phi = r <= -1.0 ? PI / 3.0 : (r >= 1.0 ? 0.0 : acos(r) / 3.0);
    double w0 = q + 2.0 * p * cos(phi);
    double w2 = q + 2.0 * p * cos(phi + 2.0 * PI / 3.0);
    return {w0, tr - (w0 + w2), w2};
}

__device__ inline void finish_eigenvector(V3<double>& v, const M3& a, double n1b, double n2b, double thresh) {
    const double eps = 2.220446049250313e-16;
    double norm = dot(v, v);
    double n1 = n1b + a.r[0].x * a.r[0].x;
    double n2 = n2b + a.r[1].y * a.r[1].y;
    double err = n1 * n2;
    if (n1 <= thresh) {
        v = {1.0, 0.0, 0.0};
    } else if (n2 <= thresh) {
        v = {0.0, 1.0, 0.0};
    } else if (norm < 4096.0 * eps * eps * err) {
        double t = fabs(a.r[0].y);
        double f = -a.r[0].x / a.r[0].y;
        if (fabs(a.r[1].y) > t) { t = fabs(a.r[1].y); f = -a.r[0].y / a.r[1].y; }
        if (fabs(a.r[1].z) > t) f = -a.r[0].z / a.r[1].z;
        double s = 1.0 / sqrt(1.0 + f * f);
        v = {s, f * s, 0.0};
    } else {
        v = v * sqrt(1.0 / norm);
    }
}

// rows = eigenvectors of w.x, w.y, w.z
__device__ inline M3 eigenvectors3(const M3& A0, V3<double> w) {
    const double eps = 2.220446049250313e-16;
    double wmax8eps = 8.0 * eps * fabs(w.x);
    double thresh = wmax8eps * wmax8eps;
    M3 a = A0;
    a.r[1].x = a.r[0].y; a.r[2].x = a.r[0].z; a.r[2].y = a.r[1].z;
    double n1 = a.r[0].y * a.r[0].y + a.r[0].z * a.r[0].z;
    double n2 = a.r[0].y * a.r[0].y + a.r[1].z * a.r[1].z;
    double base0 = a.r[0].y * a.r[1].z - a.r[0].z * a.r[1].y;
    double base1 = a.r[0].z * a.r[0].y - a.r[1].z * a.r[0].x;
    double a01sq = a.r[0].y * a.r[0].y;
    a.r[0].x -= w.x;
    a.r[1].y -= w.x;
    V3<double> v0 = {base0 + a.r[0].z * w.x, base1 + a.r[1].z * w.x, a.r[0].x * a.r[1].y - a01sq};
    finish_eigenvector(v0, a, n1, n2, thresh);
    V3<double> v1;
    double t = w.x - w.y;
    if (fabs(t) > wmax8eps) {
        a.r[0].x += t;
        a.r[1].y += t;
        v1 = {base0 + a.r[0].z * w.y, base1 + a.r[1].z * w.y, a.r[0].x * a.r[1].y - a01sq};
        finish_eigenvector(v1, a, n1, n2, thresh);
    } else {
        a.r[0].x += w.x;
        a.r[1].y += w.x;
        bool ok = false;
        for (int i = 0; i < 3 && !ok; i++) {
            m3_at(a, i, i) -= w.y;
            V3<double> ai = {m3_at(a, 0, i), m3_at(a, 1, i), m3_at(a, 2, i)};
            double ni = dot(ai, ai);
            ok = ni > thresh;
            if (ok) {
                v1 = cross(v0, ai);
                double nn = dot(v1, v1);
                ok = nn > 65536.0 * eps * eps * ni;
                if (ok) v1 = v1 * sqrt(1.0 / nn);
            }
        }
        if (!ok) {
            // fully degenerate: any unit vector orthogonal to v0
            double c[3] = {v0.x, v0.y, v0.z};
            int i = 0;
            while (i < 2 && c[i] == 0.0) i++;
            int j = (i + 1) % 3;
            double s = 1.0 / sqrt(c[i] * c[i] + c[j] * c[j]);
            double o[3] = {0.0, 0.0, 0.0};
            o[i] = c[j] * s;
            o[j] = -c[i] * s;
            v1 = {o[0], o[1], o[2]};
        }
    }
    M3 R;
    R.r[0] = v0;
    R.r[1] = v1;
    R.r[2] = cross(v0, v1);
    return R;
}

__device__ inline Q4<double> quat_from_matrix(const M3& A) {
    double a11 = A.r[0].x, a22 = A.r[1].y, a33 = A.r[2].z;
    double t[4] = {1.0 + a11 + a22 + a33, 1.0 + a11 - a22 - a33, 1.0 - a11 + a22 - a33, 1.0 - a11 - a22 + a33};
    int imax = 0;
    for (int i = 1; i < 4; i++)
        if (t[i] > t[imax]) imax = i;
    double f = 0.5 / sqrt(t[imax]);
    double o[4] = {0.0, 0.0, 0.0, 0.0};
    if (imax == 0) {
        o[1] = (A.r[1].z - A.r[2].y) * f; o[2] = (A.r[2].x - A.r[0].z) * f; o[3] = (A.r[0].y - A.r[1].x) * f;
    } else if (imax == 1) {
        o[0] = (A.r[1].z - A.r[2].y) * f; o[2] = (A.r[0].y + A.r[1].x) * f; o[3] = (A.r[0].z + A.r[2].x) * f;
    } else if (imax == 2) {
        o[0] = (A.r[2].x - A.r[0].z) * f; o[1] = (A.r[0].y + A.r[1].x) * f; o[3] = (A.r[1].z + A.r[2].y) * f;
    } else {
        o[0] = (A.r[0].y - A.r[1].x) * f; o[1] = (A.r[0].z + A.r[2].x) * f; o[2] = (A.r[1].z + A.r[2].y) * f;
    }
    o[imax] = t[imax] * f;
    return {o[0], o[1], o[2], o[3]};
}

}  // namespace rbhip
