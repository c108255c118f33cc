// CPU unit test of the shape choice for bodies of one size class
// (openmm_rigidbody_plugin_amd/csrc/rbhip_shapes.h): water-like samples with
// the frame's axis-sign variants and rounding noise far below the tolerance
// give one candidate per variant; noise above it, too many shapes or a
// first body at the origin give none.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "rbhip_shapes.h"

namespace {

int failures = 0;
void check(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAILED: %s\n", what.c_str());
        failures++;
    }
}

// a 3-atom body in its principal frame (nm), like a rigid water
const double kShape[3][3] = {{0.0, 0.00655627, 0.0}, {0.0, -0.0520320, 0.0756950}, {0.0, -0.0520320, -0.0756950}};

// ns bodies: body b takes the sign pattern signs[b % nsign] on axes 2 and 3,
// plus noise of size eps (alternating signs) on every coordinate
std::vector<float> sample(int ns, const std::vector<std::pair<int, int>>& signs, double eps) {
    const int N = 3;
    std::vector<float> s(3 * (size_t)ns * N);
    for (int b = 0; b < ns; b++) {
        const auto sg = signs[b % signs.size()];
        for (int j = 0; j < N; j++)
            for (int c = 0; c < 3; c++) {
                double v = kShape[j][c] * (c == 1 ? sg.first : c == 2 ? sg.second : 1);
                v += eps * (((b + j + c) % 2) ? 1.0 : -1.0);
                s[c * (size_t)ns * N + (size_t)b * N + j] = (float)v;
            }
    }
    return s;
}

}  // namespace

int main() {
    std::vector<int> cands;
    double tol = 0.0;

    // the frame's sign choices: four shapes; noise 1e-17 nm (FP64 rounding of
    // the frames) is far inside the tolerance
    std::vector<std::pair<int, int>> four = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    auto s = sample(32, four, 1e-17);
    check(rbhip::pick_shapes(s.data(), 32, 3, 8, cands, tol), "four sign variants share");
    check(cands == std::vector<int>({0, 1, 2, 3}), "the first body of each variant is its candidate");
    check(tol == std::ldexp((double)(float)0.0756950, -21), "tolerance is 2^-21 of the first body's size");

    // noise of half the tolerance still matches; twice the tolerance does not
    s = sample(32, {{1, 1}}, 0.25 * std::ldexp(0.0756950, -21));
    check(rbhip::pick_shapes(s.data(), 32, 3, 8, cands, tol) && cands.size() == 1, "noise within tol: one shape");
    s = sample(32, {{1, 1}}, 2.0 * std::ldexp(0.0756950, -21));
    check(rbhip::pick_shapes(s.data(), 32, 3, 8, cands, tol) && cands.size() == 2,
          "noise above tol: the two noise patterns are two shapes");

    // more shapes than allowed: nothing is shared
    std::vector<float> many(3 * 32 * 3);
    for (size_t i = 0; i < many.size(); i++) many[i] = (float)(0.01 * std::sin(1.0 + 7.0 * i));
    check(!rbhip::pick_shapes(many.data(), 32, 3, 8, cands, tol) && cands.empty(), "distinct bodies share nothing");

    // a first body at the origin has no scale to set the tolerance from
    std::vector<float> zero(3 * 4 * 3, 0.0f);
    check(!rbhip::pick_shapes(zero.data(), 4, 3, 8, cands, tol), "zero scale shares nothing");

    if (failures == 0) std::printf("Done\n");
    return failures == 0 ? 0 : 1;
}
