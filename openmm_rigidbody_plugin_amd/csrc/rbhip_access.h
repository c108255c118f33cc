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
template <class Real> __device__ __forceinline__ V3<Real> ld3(const void* base, int i) {
    const Real* p = reinterpret_cast<const Real*>(base) + 3 * (size_t)i;
    return {p[0], p[1], p[2]};
}
template <class Real> __device__ __forceinline__ void st3(void* base, int i, V3<Real> v) {
    Real* p = reinterpret_cast<Real*>(base) + 3 * (size_t)i;
    p[0] = v.x; p[1] = v.y; p[2] = v.z;
}

// position of internal atom i in FP64 (hi + lo in mixed precision)
template <class Real> __device__ __forceinline__ V3<double> load_pos(const DevSys& S, int i) {
    V3<double> p = cvt<double>(ld3<Real>(S.pos, i));
    if (sizeof(Real) == 4) {
        const float* lo = S.posLo + 3 * (size_t)i;
        p.x += (double)lo[0]; p.y += (double)lo[1]; p.z += (double)lo[2];
    }
    return p;
}
// stores an FP64 position: hi = fp32(p), lo = fp32(p - hi) in mixed precision
template <class Real> __device__ __forceinline__ void store_pos(const DevSys& S, int i, V3<double> p) {
    if (sizeof(Real) == 4) {
        float hx = (float)p.x, hy = (float)p.y, hz = (float)p.z;
        float* hi = reinterpret_cast<float*>(S.pos) + 3 * (size_t)i;
        float* lo = S.posLo + 3 * (size_t)i;
        hi[0] = hx; hi[1] = hy; hi[2] = hz;
        lo[0] = (float)(p.x - (double)hx); lo[1] = (float)(p.y - (double)hy); lo[2] = (float)(p.z - (double)hz);
    } else {
        st3<Real>(S.pos, i, cvt<Real>(p));
    }
}

}  // namespace rbhip
