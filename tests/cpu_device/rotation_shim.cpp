// Host build of the exact free-rotor rotations of rbhip_device.h for the
// tests: which = 0 runs exact_rotation (series and quadrature for short
// steps, the closed form otherwise), which = 1 the closed form alone,
// which = 2 exact_rotation as the kernels call it for fp32 state.
// in = I(3), q(4), body-frame L(3); out = q(4), L(3).
#include "rbhip_device.h"

using namespace rbhip;

extern "C" void rotate_body(int which, const double* in, double* out, double dt) {
    V3<double> I{in[0], in[1], in[2]};
    V3<double> invI{1.0 / in[0], 1.0 / in[1], 1.0 / in[2]};
    Q4<double> q{in[3], in[4], in[5], in[6]};
    V3<double> L{in[7], in[8], in[9]};
    if (which == 0)
        exact_rotation(q, L, I, invI, dt);
    else if (which == 2)
        exact_rotation<true>(q, L, I, invI, dt);
    else
        exact_rotation_closed(q, L, I, invI, dt);
    double r[7] = {q.a, q.b, q.c, q.d, L.x, L.y, L.z};
    for (int i = 0; i < 7; i++) out[i] = r[i];
}
