// Host-side choice of the shapes that the bodies of one size class share
// (no HIP dependency, so the logic is unit-tested on the CPU).
//
// Rigid bodies of one molecule type have the same body-fixed coordinates d
// up to the principal frame's axis signs and rounding far below fp32's.  From
// a sample of a class's bodies, pick_shapes picks candidate shapes greedily:
// a body whose coordinates all lie within tol of an earlier candidate's is
// that shape, any other body is a new candidate.  tol is 2^-21 of the largest
// |coordinate| of the sample's first body (four fp32 ulps at its scale).
// With more than maxShapes candidates, or a first body at the origin, the
// class shares nothing.  k_match_frames (rbhip_setup.hip) then compares every
// body of the class with the candidates by the same rule, in FP64 on the
// fp32 values.
#pragma once

#include <cmath>
#include <vector>

namespace rbhip {

// sample: three component rows of ns * N values (body b's atom j at b * N + j)
inline bool pick_shapes(const float* sample, int ns, int N, int maxShapes, std::vector<int>& cands, double& tol) {
    cands.clear();
    tol = 0.0;
    if (ns <= 0 || N <= 0) return false;
    const size_t len = (size_t)ns * N;
    auto at = [&](int body, int j, int comp) { return (double)sample[comp * len + (size_t)body * N + j]; };
    double scale = 0.0;
    for (int j = 0; j < N; j++)
        for (int comp = 0; comp < 3; comp++) scale = std::fmax(scale, std::fabs(at(0, j, comp)));
    if (!(scale > 0.0)) return false;
    tol = std::ldexp(scale, -21);
    for (int b = 0; b < ns; b++) {
        bool found = false;
        for (int cb : cands) {
            bool same = true;
            for (int j = 0; same && j < N; j++)
                for (int comp = 0; same && comp < 3; comp++) same = std::fabs(at(b, j, comp) - at(cb, j, comp)) <= tol;
            if (same) {
                found = true;
                break;
            }
        }
        if (found) continue;
        if ((int)cands.size() == maxShapes) {
            cands.clear();
            return false;
        }
        cands.push_back(b);
    }
    return true;
}

}  // namespace rbhip
