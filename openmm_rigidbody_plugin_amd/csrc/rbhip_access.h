// Typed access to the structure-of-arrays device state described by DevSys.
#pragma once

#include <hip/hip_runtime.h>

#include "rbhip_device.h"
#include "rbhip_internal.h"

namespace rbhip {

template <class Real> struct Vec4Of;
template <> struct Vec4Of<float> { using type = float4; };
template <> struct Vec4Of<double> { using type = double4; };

template <class Real> __device__ __forceinline__ Q4<Real> ld4(const void* base, int k) {
    typename Vec4Of<Real>::type v = reinterpret_cast<const typename Vec4Of<Real>::type*>(base)[k];
    return {v.x, v.y, v.z, v.w};
}
template <class Real> __device__ __forceinline__ void st4(void* base, int k, Q4<Real> q) {
    typename Vec4Of<Real>::type v;
    v.x = q.a; v.y = q.b; v.z = q.c; v.w = q.d;
    reinterpret_cast<typename Vec4Of<Real>::type*>(base)[k] = v;
}
// element i of a component-major array Real[3][stride]
template <class Real> __device__ __forceinline__ V3<Real> ld3(const void* base, int stride, int i) {
    const Real* p = reinterpret_cast<const Real*>(base) + i;
    return {p[0], p[stride], p[2 * (size_t)stride]};
}
template <class Real> __device__ __forceinline__ void st3(void* base, int stride, int i, V3<Real> v) {
    Real* p = reinterpret_cast<Real*>(base) + i;
    p[0] = v.x;
    p[stride] = v.y;
    p[2 * (size_t)stride] = v.z;
}

// element k of a packed array Real[3*n] (xyz per element)
template <class Real> __device__ __forceinline__ V3<Real> ldp3(const void* base, int k) {
    const Real* p = reinterpret_cast<const Real*>(base) + 3 * (size_t)k;
    return {p[0], p[1], p[2]};
}
template <class Real> __device__ __forceinline__ void stp3(void* base, int k, V3<Real> v) {
    Real* p = reinterpret_cast<Real*>(base) + 3 * (size_t)k;
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// position of internal atom i in FP64 (hi + lo in mixed precision)
template <class Real> __device__ __forceinline__ V3<double> load_pos(const DevSys& S, int i) {
    V3<double> p = cvt<double>(ld3<Real>(S.pos, S.atomStride, i));
    if (sizeof(Real) == 4) {
        V3<float> lo = ld3<float>(S.posLo, S.atomStride, i);
        p.x += (double)lo.x; p.y += (double)lo.y; p.z += (double)lo.z;
    }
    return p;
}
// stores an FP64 position: hi = fp32(p), lo = fp32(p - hi) in mixed precision
template <class Real> __device__ __forceinline__ void store_pos(const DevSys& S, int i, V3<double> p) {
    if (sizeof(Real) == 4) {
        V3<float> hi = {(float)p.x, (float)p.y, (float)p.z};
        V3<float> lo = {(float)(p.x - (double)hi.x), (float)(p.y - (double)hi.y), (float)(p.z - (double)hi.z)};
        st3<float>(S.pos, S.atomStride, i, hi);
        st3<float>(S.posLo, S.atomStride, i, lo);
    } else {
        st3<Real>(S.pos, S.atomStride, i, cvt<Real>(p));
    }
}

// ---------------------------------------------------------------------------
// Where the step kernels find the atom state.  Both policies give the same
// values to the same arithmetic, so a trajectory does not depend on the
// layout (given forces both layouts represent exactly).
// ---------------------------------------------------------------------------
// the library's own component-major arrays (internal atom order)
struct SoaAtoms {
    static constexpr bool adopted = false;
    template <class Real> static __device__ __forceinline__ V3<Real> force(const DevSys& S, int i) {
        return ld3<Real>(S.force, S.atomStride, i);
    }
    // FP64 position of a free atom (hi + lo in mixed precision)
    template <class Real> static __device__ __forceinline__ V3<double> pos(const DevSys& S, int i) {
        return load_pos<Real>(S, i);
    }
    template <class Real> static __device__ __forceinline__ void set_pos(const DevSys& S, int i, V3<double> p) {
        store_pos<Real>(S, i, p);
    }
    // a body atom keeps its device-precision position only: its exact value
    // is r_cm + A(q)^T d, rebuilt when positions are read back
    template <class Real> static __device__ __forceinline__ void set_body_pos(const DevSys& S, int i, V3<double> p) {
        st3<Real>(S.pos, S.atomStride, i, cvt<Real>(p));
    }
    template <class Real> static __device__ __forceinline__ V3<Real> vel(const DevSys& S, int i) {
        return ld3<Real>(S.vel, S.atomStride, i);
    }
    template <class Real> static __device__ __forceinline__ void set_vel(const DevSys& S, int i, V3<Real> v) {
        st3<Real>(S.vel, S.atomStride, i, v);
    }
};

// a caller's arrays in OpenMM's GPU layout, reached through ommSlot: posq
// (+ posqCorrection in mixed precision) as Real4, velm as double4 and the
// fixed-point force array; w components are never written
struct AdoptedAtoms {
    static constexpr bool adopted = true;
    template <class Real> static __device__ __forceinline__ V3<Real> force(const DevSys& S, int i) {
        const long long* f = S.ommForce + S.ommSlot[i];
        const size_t st = (size_t)S.ommStride;
        const double u = 1.0 / kFixedPointScale;
        return {(Real)((double)f[0] * u), (Real)((double)f[st] * u), (Real)((double)f[2 * st] * u)};
    }
    template <class Real> static __device__ __forceinline__ V3<double> pos(const DevSys& S, int i) {
        const int s = S.ommSlot[i];
        Q4<Real> p = ld4<Real>(S.ommPosq, s);
        V3<double> r = {(double)p.a, (double)p.b, (double)p.c};
        if (sizeof(Real) == 4) {
            const float* c = S.ommPosqCorr + 4 * (size_t)s;
            r.x += (double)c[0];
            r.y += (double)c[1];
            r.z += (double)c[2];
        }
        return r;
    }
    template <class Real> static __device__ __forceinline__ void set_pos(const DevSys& S, int i, V3<double> p) {
        const int s = S.ommSlot[i];
        Real* q = reinterpret_cast<Real*>(S.ommPosq) + 4 * (size_t)s;
        if (sizeof(Real) == 4) {
            float* c = S.ommPosqCorr + 4 * (size_t)s;
            float hx = (float)p.x, hy = (float)p.y, hz = (float)p.z;
            q[0] = hx; q[1] = hy; q[2] = hz;
            c[0] = (float)(p.x - (double)hx);
            c[1] = (float)(p.y - (double)hy);
            c[2] = (float)(p.z - (double)hz);
        } else {
            q[0] = (Real)p.x; q[1] = (Real)p.y; q[2] = (Real)p.z;
        }
    }
    // OpenMM's force kernels read posq (and the correction), so body atoms
    // are stored at full precision like free atoms
    template <class Real> static __device__ __forceinline__ void set_body_pos(const DevSys& S, int i, V3<double> p) {
        set_pos<Real>(S, i, p);
    }
    template <class Real> static __device__ __forceinline__ V3<Real> vel(const DevSys& S, int i) {
        const double* v = S.ommVelm + 4 * (size_t)S.ommSlot[i];
        return {(Real)v[0], (Real)v[1], (Real)v[2]};
    }
    template <class Real> static __device__ __forceinline__ void set_vel(const DevSys& S, int i, V3<Real> v) {
        double* o = S.ommVelm + 4 * (size_t)S.ommSlot[i];
        o[0] = (double)v.x; o[1] = (double)v.y; o[2] = (double)v.z;
    }
};

}  // namespace rbhip
