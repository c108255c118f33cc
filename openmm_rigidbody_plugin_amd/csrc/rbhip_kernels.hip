// Hot-path kernels of the rigid-body step for gfx950 (wave64).
//
//   k_part1  integratePart1 (reference openmmapi/src/RigidBodySystem.cpp:170-187):
//            free atoms: v += F/m dt/2, r += v dt; bodies: r_cm += p_cm dt/M
//            (FP64), free rotation (NO-SQUISH with an FP64 composite rotation,
//            or the FP64 exact solution), FP64 renormalisation of q, atom
//            positions r_j = r_cm + A(q)^T d_j stored as fp32 hi + lo.
//   k_part2  integratePart2 (RigidBodySystem.cpp:193-204): free atoms
//            v += F/m dt/2; bodies: F = sum f_j, tau = sum delta_j x f_j (FP64),
//            half-kick of p_cm and pi, atom velocities v_cm + w_s x delta_j,
//            FP64 kinetic-energy partials per tile, then the next step's first
//            half-kick (the reference applies it at the start of part1 with the
//            same F and C(q) tau, so folding it here is the same arithmetic).
//   k_fused  part2 of step n + part1 of step n+1 in one pass (interior steps
//            of RigidBodyIntegrator::step(k)).
//   k_forces the force-field slot: harmonic tethers and harmonic bonds.
//
// Every workgroup walks the tile list with a grid stride, so one launch
// covers any number of systems (the batch path) and any grid cap.
#include <hip/hip_runtime.h>

#include "rbhip_access.h"
#include "rbhip_kernels.h"

namespace rbhip {

__device__ __forceinline__ int lane_shift(int lanes) { return __ffs(lanes) - 1; }

// ---------------------------------------------------------------------------
// block reduction of up to 4 doubles per lane into slot `slot` of `out`
// (fixed order: wave shuffles, then waves in index order)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <int NV>
__device__ void block_reduce_store(double (&v)[NV], double* out, int slot) {
    __shared__ double red[8][4];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < NV; c++) {
        double s = wave_sum(v[c]);
        if (lane == 0) red[wave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int c = 0; c < NV; c++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += red[w][c];
            out[4 * (size_t)slot + c] = s;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// part 1
// ---------------------------------------------------------------------------
template <class Real>
__device__ __forceinline__ void part1_free(const DevSys& S, const Tile& tile) {
    int i = threadIdx.x;
    if (i >= tile.count) return;
    int a = tile.first + i;
    Real halfDt = (Real)(0.5 * S.dt);
    V3<Real> v = ld3<Real>(S.vel, S.atomStride, a);
    V3<Real> f = ld3<Real>(S.force, S.atomStride, a);
    Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
    v = v + f * invm * halfDt;
    V3<double> r = load_pos<Real>(S, a);
    r = r + cvt<double>(v) * S.dt;
    st3<Real>(S.vel, S.atomStride, a, v);
    store_pos<Real>(S, a, r);
}

template <class Real, int NC>
__device__ __forceinline__ void part1_atoms(const DevSys& S, const Rot<Real>& R, V3<double> rc, int atom0,
                                            int N, int sub, int lanes) {
    const int dbase = atom0 - S.numFree;
    if constexpr (NC > 0) {
        V3<Real> dj[NC > 0 ? NC : 1];
#pragma unroll
        for (int j = 0; j < NC; j++) dj[j] = ld3<Real>(S.d, S.dStride, dbase + j);
#pragma unroll
        for (int j = 0; j < NC; j++) store_pos<Real>(S, atom0 + j, rc + cvt<double>(apply(R, dj[j])));
    } else {
        for (int j = sub; j < N; j += lanes) {
            V3<Real> dj = ld3<Real>(S.d, S.dStride, dbase + j);
            store_pos<Real>(S, atom0 + j, rc + cvt<double>(apply(R, dj)));
        }
    }
}

template <class Real, bool EXACT>
__device__ __forceinline__ void part1_bodies(const DevSys& S, const Tile& tile) {
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    if (b >= tile.count) return;
    const int k = tile.first + b;
    const int N = tile.N;
    const int atom0 = tile.atomFirst + b * N;
    Q4<Real> q = ld4<Real>(S.q, k);
    Q4<Real> pi = ld4<Real>(S.pi, k);
    Q4<Real> pm = ld4<Real>(S.pm, k);
    Q4<Real> iI = ld4<Real>(S.iI, k);
    const int st = S.bodyStride;
    V3<double> rc = {S.rcm[k], S.rcm[k + st], S.rcm[k + 2 * st]};
    // r_cm += p_cm (dt / M)
    double idt = (double)pm.d * S.dt;
    rc = rc + V3<double>{(double)pm.a, (double)pm.b, (double)pm.c} * idt;
    V3<Real> invI = {iI.a, iI.b, iI.c};
    if (EXACT) {
        Q4<double> qd = cvt<double>(q), pd = cvt<double>(pi);
        V3<double> invId = cvt<double>(invI);
        if (invI.z == (Real)0) {
            linear_rotor_exact(qd, pd, invId, S.dt);
        } else {
            V3<double> I = {1.0 / invId.x, 1.0 / invId.y, 1.0 / invId.z};
            exact_rotation(qd, pd, I, invId, S.dt);
        }
        q = cvt<Real>(qd);
        pi = cvt<Real>(pd);
    } else {
        no_squish<Real>(q, pi, invI, S.dt, S.rotationMode);
    }
    if (sub == 0) {
        st4<Real>(S.q, k, q);
        st4<Real>(S.pi, k, pi);
        S.rcm[k] = rc.x;
        S.rcm[k + st] = rc.y;
        S.rcm[k + 2 * st] = rc.z;
    }
    Rot<Real> R = rotT(q);
    if (tile.lanes == 1) {
        switch (N) {
        case 2: part1_atoms<Real, 2>(S, R, rc, atom0, N, 0, 1); break;
        case 3: part1_atoms<Real, 3>(S, R, rc, atom0, N, 0, 1); break;
        case 4: part1_atoms<Real, 4>(S, R, rc, atom0, N, 0, 1); break;
        case 6: part1_atoms<Real, 6>(S, R, rc, atom0, N, 0, 1); break;
        default: part1_atoms<Real, 0>(S, R, rc, atom0, N, 0, 1); break;
        }
    } else {
        part1_atoms<Real, 0>(S, R, rc, atom0, N, sub, tile.lanes);
    }
}

template <class Real, bool EXACT>
__global__ __launch_bounds__(EXACT ? 256 : 512) void k_part1(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                 int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        if (tile.kind == TILE_FREE)
            part1_free<Real>(S, tile);
        else
            part1_bodies<Real, EXACT>(S, tile);
    }
}

// ---------------------------------------------------------------------------
// part 2
// ---------------------------------------------------------------------------
template <class Real>
__device__ __forceinline__ void part2_free(const DevSys& S, const Tile& tile, double (&ke)[2]) {
    int i = threadIdx.x;
    if (i >= tile.count) return;
    int a = tile.first + i;
    Real halfDt = (Real)(0.5 * S.dt);
    V3<Real> v = ld3<Real>(S.vel, S.atomStride, a);
    V3<Real> f = ld3<Real>(S.force, S.atomStride, a);
    Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
    v = v + f * invm * halfDt;
    st3<Real>(S.vel, S.atomStride, a, v);
    V3<double> vd = cvt<double>(v);
    ke[0] = dot(vd, vd) / (double)invm;
}

template <class Real, int NC>
__device__ __forceinline__ void part2_body_sums(const DevSys& S, const Rot<Real>& R, int atom0, int N, int sub,
                                                int lanes, V3<double>& F, V3<double>& tau) {
    const int dbase = atom0 - S.numFree;
    if constexpr (NC > 0) {
        V3<Real> dj[NC > 0 ? NC : 1], fj[NC > 0 ? NC : 1];
#pragma unroll
        for (int j = 0; j < NC; j++) {
            dj[j] = ld3<Real>(S.d, S.dStride, dbase + j);
            fj[j] = ld3<Real>(S.force, S.atomStride, atom0 + j);
        }
#pragma unroll
        for (int j = 0; j < NC; j++) {
            V3<double> dl = cvt<double>(apply(R, dj[j]));
            V3<double> f = cvt<double>(fj[j]);
            F = F + f;
            tau = tau + cross(dl, f);
        }
    } else {
        for (int j = sub; j < N; j += lanes) {
            V3<double> dl = cvt<double>(apply(R, ld3<Real>(S.d, S.dStride, dbase + j)));
            V3<double> f = cvt<double>(ld3<Real>(S.force, S.atomStride, atom0 + j));
            F = F + f;
            tau = tau + cross(dl, f);
        }
    }
}

template <class Real, int NC>
__device__ __forceinline__ void part2_velocities(const DevSys& S, const Rot<Real>& R, V3<Real> vcm, V3<Real> ws,
                                                 int atom0, int N, int sub, int lanes) {
    const int dbase = atom0 - S.numFree;
    if constexpr (NC > 0) {
#pragma unroll
        for (int j = 0; j < NC; j++) {
            V3<Real> dl = apply(R, ld3<Real>(S.d, S.dStride, dbase + j));
            st3<Real>(S.vel, S.atomStride, atom0 + j, vcm + cross(ws, dl));
        }
    } else {
        for (int j = sub; j < N; j += lanes) {
            V3<Real> dl = apply(R, ld3<Real>(S.d, S.dStride, dbase + j));
            st3<Real>(S.vel, S.atomStride, atom0 + j, vcm + cross(ws, dl));
        }
    }
}

__device__ __forceinline__ double xor_sum(double v, int lanes) {
    for (int m = 1; m < lanes; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <class Real>
__device__ __forceinline__ void part2_bodies(const DevSys& S, const Tile& tile, double (&ke)[2]) {
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    const bool active = b < tile.count;
    const int k = tile.first + (active ? b : 0);
    const int N = tile.N;
    const int atom0 = tile.atomFirst + (active ? b : 0) * N;
    Q4<Real> q = ld4<Real>(S.q, k);
    Q4<Real> pi = ld4<Real>(S.pi, k);
    Q4<Real> pm = ld4<Real>(S.pm, k);
    Q4<Real> iI = ld4<Real>(S.iI, k);
    Rot<Real> R = rotT(q);
    V3<double> F = {0.0, 0.0, 0.0}, tau = {0.0, 0.0, 0.0};
    if (tile.lanes == 1) {
        if (active) {
            switch (N) {
            case 2: part2_body_sums<Real, 2>(S, R, atom0, N, 0, 1, F, tau); break;
            case 3: part2_body_sums<Real, 3>(S, R, atom0, N, 0, 1, F, tau); break;
            case 4: part2_body_sums<Real, 4>(S, R, atom0, N, 0, 1, F, tau); break;
            case 6: part2_body_sums<Real, 6>(S, R, atom0, N, 0, 1, F, tau); break;
            default: part2_body_sums<Real, 0>(S, R, atom0, N, 0, 1, F, tau); break;
            }
        }
    } else {
        if (active) part2_body_sums<Real, 0>(S, R, atom0, N, sub, tile.lanes, F, tau);
        F = {xor_sum(F.x, tile.lanes), xor_sum(F.y, tile.lanes), xor_sum(F.z, tile.lanes)};
        tau = {xor_sum(tau.x, tile.lanes), xor_sum(tau.y, tile.lanes), xor_sum(tau.z, tile.lanes)};
    }
    if (!active) return;
    const double dt = S.dt, halfDt = 0.5 * dt;
    Q4<double> qd = cvt<double>(q);
    Q4<double> torque = Cmul(qd, tau);
    V3<double> pcm = V3<double>{(double)pm.a, (double)pm.b, (double)pm.c} + F * halfDt;
    Q4<double> pid = cvt<double>(pi) + torque * dt;
    double invM = (double)pm.d;
    V3<double> L = Btmul(qd, pid) * 0.5;
    V3<double> w = hadamard(cvt<double>(V3<Real>{iI.a, iI.b, iI.c}), L);
    V3<double> vcm = pcm * invM;
    V3<Real> ws = apply(R, cvt<Real>(w));
    if (tile.lanes == 1) {
        switch (N) {
        case 2: part2_velocities<Real, 2>(S, R, cvt<Real>(vcm), ws, atom0, N, 0, 1); break;
        case 3: part2_velocities<Real, 3>(S, R, cvt<Real>(vcm), ws, atom0, N, 0, 1); break;
        case 4: part2_velocities<Real, 4>(S, R, cvt<Real>(vcm), ws, atom0, N, 0, 1); break;
        case 6: part2_velocities<Real, 6>(S, R, cvt<Real>(vcm), ws, atom0, N, 0, 1); break;
        default: part2_velocities<Real, 0>(S, R, cvt<Real>(vcm), ws, atom0, N, 0, 1); break;
        }
    } else {
        part2_velocities<Real, 0>(S, R, cvt<Real>(vcm), ws, atom0, N, sub, tile.lanes);
    }
    if (sub == 0) {
        ke[0] = dot(pcm, vcm);
        ke[1] = dot(L, w);
        // first half-kick of the next step (same F and torque)
        pcm = pcm + F * halfDt;
        pid = pid + torque * dt;
        st4<Real>(S.pm, k, Q4<Real>{(Real)pcm.x, (Real)pcm.y, (Real)pcm.z, pm.d});
        st4<Real>(S.pi, k, cvt<Real>(pid));
    }
}

template <class Real>
__global__ __launch_bounds__(512) void k_part2(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                 int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        double ke[2] = {0.0, 0.0};
        if (tile.kind == TILE_FREE)
            part2_free<Real>(S, tile, ke);
        else
            part2_bodies<Real>(S, tile, ke);
        block_reduce_store<2>(ke, S.kePart, tile.slot);
    }
}

// ---------------------------------------------------------------------------
// fused step boundary: part2 of step n followed by part1 of step n+1 for the
// same body, one pass over its data.  Inside RigidBodyIntegrator::step(k) the
// atom velocities and kinetic energies of the interior steps are read by
// nothing (forces depend on positions only), so this pass writes neither;
// the last step of the call ends with a plain part2.  The body state is
// rounded to storage precision at the same points as part2 + part1, so the
// trajectory is bit-identical to the two-pass path.
// ---------------------------------------------------------------------------
template <class Real>
__device__ __forceinline__ void fused_free(const DevSys& S, const Tile& tile) {
    int i = threadIdx.x;
    if (i >= tile.count) return;
    int a = tile.first + i;
    Real halfDt = (Real)(0.5 * S.dt);
    V3<Real> v = ld3<Real>(S.vel, S.atomStride, a);
    V3<Real> f = ld3<Real>(S.force, S.atomStride, a);
    Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
    v = v + f * invm * halfDt;     // end of step n
    v = v + f * invm * halfDt;     // start of step n+1
    V3<double> r = load_pos<Real>(S, a);
    r = r + cvt<double>(v) * S.dt;
    st3<Real>(S.vel, S.atomStride, a, v);
    store_pos<Real>(S, a, r);
}

template <class Real, bool EXACT>
__device__ __forceinline__ void fused_bodies(const DevSys& S, const Tile& tile) {
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    const bool active = b < tile.count;
    const int k = tile.first + (active ? b : 0);
    const int N = tile.N;
    const int atom0 = tile.atomFirst + (active ? b : 0) * N;
    Q4<Real> q = ld4<Real>(S.q, k);
    Q4<Real> pi = ld4<Real>(S.pi, k);
    Q4<Real> pm = ld4<Real>(S.pm, k);
    Q4<Real> iI = ld4<Real>(S.iI, k);
    const int st = S.bodyStride;
    V3<double> rc = {S.rcm[k], S.rcm[k + st], S.rcm[k + 2 * st]};
    // part2 of step n: force and torque, half-kick, then the first half-kick of step n+1
    Rot<Real> R = rotT(q);
    V3<double> F = {0.0, 0.0, 0.0}, tau = {0.0, 0.0, 0.0};
    if (tile.lanes == 1) {
        if (active) {
            switch (N) {
            case 2: part2_body_sums<Real, 2>(S, R, atom0, N, 0, 1, F, tau); break;
            case 3: part2_body_sums<Real, 3>(S, R, atom0, N, 0, 1, F, tau); break;
            case 4: part2_body_sums<Real, 4>(S, R, atom0, N, 0, 1, F, tau); break;
            case 6: part2_body_sums<Real, 6>(S, R, atom0, N, 0, 1, F, tau); break;
            default: part2_body_sums<Real, 0>(S, R, atom0, N, 0, 1, F, tau); break;
            }
        }
    } else {
        if (active) part2_body_sums<Real, 0>(S, R, atom0, N, sub, tile.lanes, F, tau);
        F = {xor_sum(F.x, tile.lanes), xor_sum(F.y, tile.lanes), xor_sum(F.z, tile.lanes)};
        tau = {xor_sum(tau.x, tile.lanes), xor_sum(tau.y, tile.lanes), xor_sum(tau.z, tile.lanes)};
    }
    if (!active) return;
    const double dt = S.dt, halfDt = 0.5 * dt;
    Q4<double> torque = Cmul(cvt<double>(q), tau);
    V3<double> pcm = V3<double>{(double)pm.a, (double)pm.b, (double)pm.c} + F * halfDt;
    Q4<double> pid = cvt<double>(pi) + torque * dt;
    pcm = pcm + F * halfDt;
    pid = pid + torque * dt;
    pm = Q4<Real>{(Real)pcm.x, (Real)pcm.y, (Real)pcm.z, pm.d};
    pi = cvt<Real>(pid);
    // part1 of step n+1: drift, rotation, atom positions
    double idt = (double)pm.d * dt;
    rc = rc + V3<double>{(double)pm.a, (double)pm.b, (double)pm.c} * idt;
    V3<Real> invI = {iI.a, iI.b, iI.c};
    if (EXACT) {
        Q4<double> qd = cvt<double>(q), pd = cvt<double>(pi);
        V3<double> invId = cvt<double>(invI);
        if (invI.z == (Real)0) {
            linear_rotor_exact(qd, pd, invId, dt);
        } else {
            V3<double> I = {1.0 / invId.x, 1.0 / invId.y, 1.0 / invId.z};
            exact_rotation(qd, pd, I, invId, dt);
        }
        q = cvt<Real>(qd);
        pi = cvt<Real>(pd);
    } else {
        no_squish<Real>(q, pi, invI, dt, S.rotationMode);
    }
    if (sub == 0) {
        st4<Real>(S.q, k, q);
        st4<Real>(S.pi, k, pi);
        st4<Real>(S.pm, k, pm);
        S.rcm[k] = rc.x;
        S.rcm[k + st] = rc.y;
        S.rcm[k + 2 * st] = rc.z;
    }
    Rot<Real> R1 = rotT(q);
    if (tile.lanes == 1) {
        switch (N) {
        case 2: part1_atoms<Real, 2>(S, R1, rc, atom0, N, 0, 1); break;
        case 3: part1_atoms<Real, 3>(S, R1, rc, atom0, N, 0, 1); break;
        case 4: part1_atoms<Real, 4>(S, R1, rc, atom0, N, 0, 1); break;
        case 6: part1_atoms<Real, 6>(S, R1, rc, atom0, N, 0, 1); break;
        default: part1_atoms<Real, 0>(S, R1, rc, atom0, N, 0, 1); break;
        }
    } else {
        part1_atoms<Real, 0>(S, R1, rc, atom0, N, sub, tile.lanes);
    }
}

template <class Real, bool EXACT>
__global__ __launch_bounds__(EXACT ? 256 : 512) void k_fused(const DevSys* __restrict__ systems,
                                                              const Tile* __restrict__ tiles, int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        if (tile.kind == TILE_FREE)
            fused_free<Real>(S, tile);
        else
            fused_bodies<Real, EXACT>(S, tile);
    }
}

// ---------------------------------------------------------------------------
// force field: harmonic tether + harmonic bonds, one lane per atom
// ---------------------------------------------------------------------------
template <class Real>
__global__ __launch_bounds__(512) void k_forces(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                  int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        double e[1] = {0.0};
        int i = threadIdx.x;
        if (i < tile.count) {
            int a = tile.first + i;
            V3<Real> f = {(Real)0, (Real)0, (Real)0};
            if (S.tetherScale != 0.0) {
                Real k = reinterpret_cast<const Real*>(S.tetherK)[a];
                V3<Real> dx;
                if (sizeof(Real) == 4) {
                    V3<Real> ph = ld3<Real>(S.pos, S.atomStride, a), rh = ld3<Real>(S.r0, S.atomStride, a);
                    V3<float> pl = ld3<float>(S.posLo, S.atomStride, a), rl = ld3<float>(S.r0Lo, S.atomStride, a);
                    dx = {(ph.x - rh.x) + (Real)(pl.x - rl.x), (ph.y - rh.y) + (Real)(pl.y - rl.y),
                          (ph.z - rh.z) + (Real)(pl.z - rl.z)};
                } else {
                    dx = ld3<Real>(S.pos, S.atomStride, a) - ld3<Real>(S.r0, S.atomStride, a);
                }
                f = dx * (-k);
                V3<double> dd = cvt<double>(dx);
                e[0] += 0.5 * (double)k * dot(dd, dd);
            }
            if (S.bondStart) {
                int e0 = S.bondStart[a], e1 = S.bondStart[a + 1];
                if (e0 < e1) {
                    V3<double> pa = load_pos<Real>(S, a);
                    V3<double> fd = cvt<double>(f);
                    for (int b = e0; b < e1; b++) {
                        V3<double> dr = load_pos<Real>(S, S.bondOther[b]) - pa;
                        double r = sqrt(dot(dr, dr));
                        double ext = r - S.bondLen[b];
                        fd = fd + dr * (S.bondK[b] * ext / r);
                        e[0] += 0.25 * S.bondK[b] * ext * ext;
                    }
                    f = cvt<Real>(fd);
                }
            }
            st3<Real>(S.force, S.atomStride, a, f);
        }
        block_reduce_store<1>(e, S.pePart, tile.slot);
    }
}

// deterministic sum of `nslots` 4-double partials of each system; one block per system
__global__ __launch_bounds__(256) void k_reduce(const DevSys* __restrict__ systems, const int* __restrict__ nslots,
                                                 int which, double* __restrict__ out) {
    const DevSys& S = systems[blockIdx.x];
    const double* part = which == 0 ? S.kePart : S.pePart;
    int n = nslots[blockIdx.x];
    __shared__ double red[256][4];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = threadIdx.x; s < n; s += blockDim.x)
        for (int c = 0; c < 4; c++) acc[c] += part[4 * (size_t)s + c];
    for (int c = 0; c < 4; c++) red[threadIdx.x][c] = acc[c];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; c++) red[threadIdx.x][c] += red[threadIdx.x + w][c];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 4; c++) out[4 * (size_t)blockIdx.x + c] = red[0][c];
}

__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int grid_for(int numTiles, int maxBlocks) {
    if (numTiles <= 0) return 0;
    return maxBlocks > 0 && maxBlocks < numTiles ? maxBlocks : numTiles;
}

hipError_t launch_part1(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                        int block, int maxBlocks, hipStream_t stream) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) {
        if (rotationMode == 0) k_part1<float, true><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
        else k_part1<float, false><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    } else {
        if (rotationMode == 0) k_part1<double, true><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
        else k_part1<double, false><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    }
    return hipGetLastError();
}

hipError_t launch_fused(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                        int block, int maxBlocks, hipStream_t stream) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) {
        if (rotationMode == 0) k_fused<float, true><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
        else k_fused<float, false><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    } else {
        if (rotationMode == 0) k_fused<double, true><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
        else k_fused<double, false><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    }
    return hipGetLastError();
}

hipError_t launch_part2(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                        int maxBlocks, hipStream_t stream) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) k_part2<float><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    else k_part2<double><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    return hipGetLastError();
}

hipError_t launch_forces(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                         int maxBlocks, hipStream_t stream) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) k_forces<float><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    else k_forces<double><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    return hipGetLastError();
}

hipError_t launch_reduce(const DevSys* systems, const int* nslots, int numSystems, int which, double* out,
                         hipStream_t stream) {
    if (numSystems <= 0) return hipSuccess;
    k_reduce<<<numSystems, 256, 0, stream>>>(systems, nslots, which, out);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream) {
    size_t n = bytes / sizeof(float4);
    int grid = (int)((n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192);
    if (grid < 1) grid = 1;
    k_copy<<<grid, 256, 0, stream>>>(reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n);
    return hipGetLastError();
}

}  // namespace rbhip
