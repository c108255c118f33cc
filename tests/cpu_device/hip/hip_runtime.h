// Stand-in for the HIP runtime header that lets the device math of
// rbhip_device.h (plain FP64 arithmetic and libm calls, no intrinsics)
// compile as host code, so that tests can check it against an independent
// integration without a GPU.
#pragma once
#include <cmath>
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
