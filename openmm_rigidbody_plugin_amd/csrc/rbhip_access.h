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

}  // namespace rbhip
