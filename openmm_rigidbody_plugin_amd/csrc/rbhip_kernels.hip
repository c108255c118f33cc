// Hot-path kernels of the rigid-body step for gfx950 (wave64).
//
//   k_part1  integratePart1 (reference openmmapi/src/RigidBodySystem.cpp:170-187):
//            free atoms: v += F/m dt/2, r += v dt; bodies: r_cm += p_cm dt/M
//            (FP64), free rotation (NO-SQUISH with an FP64-accurate composite
//            rotation, or the FP64 exact solution), FP64 renormalisation of q,
//            atom positions r_j = r_cm + A(q)^T d_j.
//   k_part2  integratePart2 (RigidBodySystem.cpp:193-204): free atoms
//            v += F/m dt/2; bodies: F = sum f_j, tau = sum delta_j x f_j (FP64),
//            half-kick of p_cm and L, atom velocities v_cm + w_s x delta_j,
//            FP64 kinetic-energy partials per tile, then the next step's first
//            half-kick (the reference applies it at the start of part1 with the
//            same F and torque, so folding it here is the same arithmetic).
//   k_fused  part2 of step n + part1 of step n+1 in one pass (interior steps
//            of RigidBodyIntegrator::step(k)).
//   k_forces the force-field slot: harmonic tethers and harmonic bonds.
//
// Body state per body: q (Real4), body-frame L and p_cm (3 Reals each), iI =
// (1/I, 1/M), r_cm (FP64).  The reference's quaternion momentum is
// pi = 2 B(q) L, its torque kick pi += C(q) tau dt is L += A(q) tau dt / 2.
// Every workgroup walks the tile list with a grid stride, so one launch
// covers any number of systems (the batch path) and any grid cap.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

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

__device__ __forceinline__ double xor_sum(double v, int lanes) {
    for (int m = 1; m < lanes; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// ---------------------------------------------------------------------------
// body state and the per-atom loops
// ---------------------------------------------------------------------------
template <class Real> struct Body {
    Q4<Real> q;
    V3<Real> L, p, invI;
    Real invM;
};

// body k's state, with 1/I and 1/M read at body ks (the body whose inertia it
// shares, or itself)
template <class Real> __device__ __forceinline__ Body<Real> load_body(const DevSys& S, int k, int ks) {
    Q4<Real> iI = ld4<Real>(S.iI, ks);
    return Body<Real>{ld4<Real>(S.q, k), ldp3<Real>(S.L, k), ldp3<Real>(S.pcm, k), V3<Real>{iI.a, iI.b, iI.c}, iI.d};
}

// atoms per lane a strided per-body loop loads before it uses any (bodies of
// more than 6 atoms, several lanes each)
constexpr int STRIDED_BATCH = 8;

// F = sum f_j and tau = sum delta_j x f_j in FP64, delta_j = A(q)^T d_j
// Where body k (first atom atom0, N atoms) reads its body-fixed coordinates
// and inertia: its own, or, in a tile of bodies that share them, those of
// body DevSys::dOff[k] of the same size class, whose rows lie (k - ks) N rows
// before its own.  ks: the body whose 1/I and 1/M it reads; row: its first
// row of d.
struct Shape { int ks, row; };
__device__ __forceinline__ Shape body_shape(const DevSys& S, const Tile& tile, int k, int atom0, int N) {
    const int own = atom0 - S.numFree;
    if (!tile.sharedD) return {k, own};
    const int ks = S.dOff[k];
    return {ks, own - (k - ks) * N};
}

template <class Real, int NC, class IO>
__device__ __forceinline__ void body_sums(const DevSys& S, const Rot<Real>& R, int atom0, int dbase, int N, int sub,
                                          int lanes, V3<double>& F, V3<double>& tau) {
    if constexpr (NC > 0) {
        V3<Real> dj[NC], fj[NC];
#pragma unroll
        for (int j = 0; j < NC; j++) {
            dj[j] = ld3<Real>(S.d, S.dStride, dbase + j);
            fj[j] = IO::template force<Real>(S, atom0 + j);
        }
#pragma unroll
        for (int j = 0; j < NC; j++) {
            V3<double> dl = cvt<double>(apply(R, dj[j]));
            V3<double> f = cvt<double>(fj[j]);
            F = F + f;
            tau = tau + cross(dl, f);
        }
    } else {
        // atoms sub, sub + lanes, ...: up to STRIDED_BATCH per lane loaded
        // before any is summed (one round trip for the bodies the default
        // lanes give), summed in the same order
        constexpr int U = STRIDED_BATCH;
        for (int j0 = sub; j0 < N; j0 += U * lanes) {
            V3<Real> dj[U], fj[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int j = j0 + u * lanes;
                if (j < N) {
                    dj[u] = ld3<Real>(S.d, S.dStride, dbase + j);
                    fj[u] = IO::template force<Real>(S, atom0 + j);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (j0 + u * lanes < N) {
                    V3<double> dl = cvt<double>(apply(R, dj[u]));
                    V3<double> f = cvt<double>(fj[u]);
                    F = F + f;
                    tau = tau + cross(dl, f);
                }
            }
        }
    }
}

template <class Real, class IO = SoaAtoms>
__device__ __forceinline__ void body_sums_any(const DevSys& S, const Rot<Real>& R, int atom0, int dbase, int N,
                                              int sub, int lanes, bool active, V3<double>& F, V3<double>& tau) {
    F = {0.0, 0.0, 0.0};
    tau = {0.0, 0.0, 0.0};
    if (lanes == 1) {
        if (active) {
            switch (N) {
            case 2: body_sums<Real, 2, IO>(S, R, atom0, dbase, N, 0, 1, F, tau); break;
            case 3: body_sums<Real, 3, IO>(S, R, atom0, dbase, N, 0, 1, F, tau); break;
            case 4: body_sums<Real, 4, IO>(S, R, atom0, dbase, N, 0, 1, F, tau); break;
            case 6: body_sums<Real, 6, IO>(S, R, atom0, dbase, N, 0, 1, F, tau); break;
            default: body_sums<Real, 0, IO>(S, R, atom0, dbase, N, 0, 1, F, tau); break;
            }
        }
    } else {
        if (active) body_sums<Real, 0, IO>(S, R, atom0, dbase, N, sub, lanes, F, tau);
        F = {xor_sum(F.x, lanes), xor_sum(F.y, lanes), xor_sum(F.z, lanes)};
        tau = {xor_sum(tau.x, lanes), xor_sum(tau.y, lanes), xor_sum(tau.z, lanes)};
    }
}

// atom positions r_cm + A(q)^T d_j: fp32 in mixed precision (the force field's
// view; the exact value is rebuilt from the body state on download)
template <class Real, int NC, class IO>
__device__ __forceinline__ void body_positions(const DevSys& S, const Rot<Real>& R, V3<double> rc, int atom0,
                                               int dbase, int N, int sub, int lanes) {
    if constexpr (NC > 0) {
        V3<Real> dj[NC];
#pragma unroll
        for (int j = 0; j < NC; j++) dj[j] = ld3<Real>(S.d, S.dStride, dbase + j);
#pragma unroll
        for (int j = 0; j < NC; j++) IO::template set_body_pos<Real>(S, atom0 + j, rc + cvt<double>(apply(R, dj[j])));
    } else {
        // loads of a batch first, then its stores (one round trip per batch)
        constexpr int U = STRIDED_BATCH;
        for (int j0 = sub; j0 < N; j0 += U * lanes) {
            V3<Real> dj[U];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j0 + u * lanes < N) dj[u] = ld3<Real>(S.d, S.dStride, dbase + j0 + u * lanes);
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j0 + u * lanes < N)
                    IO::template set_body_pos<Real>(S, atom0 + j0 + u * lanes, rc + cvt<double>(apply(R, dj[u])));
        }
    }
}

template <class Real, class IO>
__device__ __forceinline__ void body_positions_any(const DevSys& S, const Rot<Real>& R, V3<double> rc, int atom0,
                                                   int dbase, int N, int sub, int lanes) {
    if (lanes == 1) {
        switch (N) {
        case 2: body_positions<Real, 2, IO>(S, R, rc, atom0, dbase, N, 0, 1); break;
        case 3: body_positions<Real, 3, IO>(S, R, rc, atom0, dbase, N, 0, 1); break;
        case 4: body_positions<Real, 4, IO>(S, R, rc, atom0, dbase, N, 0, 1); break;
        case 6: body_positions<Real, 6, IO>(S, R, rc, atom0, dbase, N, 0, 1); break;
        default: body_positions<Real, 0, IO>(S, R, rc, atom0, dbase, N, 0, 1); break;
        }
    } else {
        body_positions<Real, 0, IO>(S, R, rc, atom0, dbase, N, sub, lanes);
    }
}

template <class Real, int NC, class IO>
__device__ __forceinline__ void body_velocities(const DevSys& S, const Rot<Real>& R, V3<Real> vcm, V3<Real> ws,
                                                int atom0, int dbase, int N, int sub, int lanes) {
    if constexpr (NC > 0) {
#pragma unroll
        for (int j = 0; j < NC; j++) {
            V3<Real> dl = apply(R, ld3<Real>(S.d, S.dStride, dbase + j));
            IO::template set_vel<Real>(S, atom0 + j, vcm + cross(ws, dl));
        }
    } else {
        constexpr int U = STRIDED_BATCH;
        for (int j0 = sub; j0 < N; j0 += U * lanes) {
            V3<Real> dj[U];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j0 + u * lanes < N) dj[u] = ld3<Real>(S.d, S.dStride, dbase + j0 + u * lanes);
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j0 + u * lanes < N)
                    IO::template set_vel<Real>(S, atom0 + j0 + u * lanes, vcm + cross(ws, apply(R, dj[u])));
        }
    }
}

// A(q) x from the rows of A(q)^T
template <class Real> __device__ __forceinline__ V3<double> to_body(const Rot<Real>& R, V3<double> x) {
    return cvt<double>(R.r0) * x.x + cvt<double>(R.r1) * x.y + cvt<double>(R.r2) * x.z;
}

// body-frame torque; a linear body (1/I2 = 0) has no moment about its axis
// and keeps L2 = 0 exactly (the axial torque of collinear atoms is rounding)
template <class Real> __device__ __forceinline__ V3<double> axial_torque(V3<double> tb, V3<Real> invI) {
    if (invI.z == (Real)0) tb.z = 0.0;
    return tb;
}

// free rotation of (q, L) over h: the FP64 exact solution or NO-SQUISH(mode)
template <class Real, bool EXACT>
__device__ __forceinline__ void free_rotation(Q4<Real>& q, V3<Real>& L, V3<Real> invI, double h, int mode) {
    if (EXACT) {
        Q4<double> qd = cvt<double>(q);
        V3<double> Ld = cvt<double>(L), invId = cvt<double>(invI);
        if (invI.z == (Real)0) {
            linear_rotor_exact(qd, Ld, invId, h);
        } else {
            V3<double> I = {1.0 / invId.x, 1.0 / invId.y, 1.0 / invId.z};
            exact_rotation<sizeof(Real) == 4>(qd, Ld, I, invId, h);
        }
        q = cvt<Real>(qd);
        L = cvt<Real>(Ld);
    } else {
        no_squish<Real>(q, L, invI, h, mode);
    }
}

// r_cm += p_cm (dt / M), then the free rotation of (q, L)
template <class Real, bool EXACT>
__device__ __forceinline__ void drift_and_rotate(const DevSys& S, Body<Real>& B, V3<double>& rc) {
    double idt = (double)B.invM * S.dt;
    rc = rc + cvt<double>(B.p) * idt;
    free_rotation<Real, EXACT>(B.q, B.L, B.invI, S.dt, S.rotationMode);
}

template <class Real> __device__ __forceinline__ void store_rcm(const DevSys& S, int k, V3<double> rc) {
    const int st = S.bodyStride;
    S.rcm[k] = rc.x;
    S.rcm[k + st] = rc.y;
    S.rcm[k + 2 * st] = rc.z;
}
template <class Real> __device__ __forceinline__ V3<double> load_rcm(const DevSys& S, int k) {
    const int st = S.bodyStride;
    return V3<double>{S.rcm[k], S.rcm[k + st], S.rcm[k + 2 * st]};
}

// ---------------------------------------------------------------------------
// part 1
// ---------------------------------------------------------------------------
// free-atom tiles hold up to FREE_PER_THREAD atoms per thread (strided by
// the workgroup size, so each pass over the tile is coalesced)
template <class Real, class IO>
__device__ __forceinline__ void part1_free(const DevSys& S, const Tile& tile) {
    Real halfDt = (Real)(0.5 * S.dt);
    for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
        int a = tile.first + i;
        V3<Real> v = IO::template vel<Real>(S, a);
        V3<Real> f = IO::template force<Real>(S, a);
        Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
        v = v + f * invm * halfDt;
        V3<double> r = IO::template pos<Real>(S, a);
        if constexpr (IO::adopted) {
            if (S.posDelta) {
                // the caller's constrained displacement; the correction to
                // the unconstrained drift waits for part 2
                const double* pd = S.posDelta + 4 * (size_t)S.ommSlot[a];
                V3<double> ru = r + cvt<double>(v) * S.dt;
                r = r + V3<double>{pd[0], pd[1], pd[2]};
                st3<Real>(S.freeCorr, S.numFree, a, cvt<Real>(r - ru));
                IO::template set_vel<Real>(S, a, v);
                IO::template set_pos<Real>(S, a, r);
                continue;
            }
        }
        r = r + cvt<double>(v) * S.dt;
        IO::template set_vel<Real>(S, a, v);
        IO::template set_pos<Real>(S, a, r);
    }
}

// The free atoms' drift displacements (v + F (1/m) dt/2) dt, as part 1
// computes them, into a caller's posDelta (double4 per slot, w = 0) for its
// constraint solver (the CUDA platform's freeAtomsDelta,
// rigidbodyintegrator.cu:270-279); adopted arrays only.
template <class Real>
__global__ void k_free_delta(const DevSys* __restrict__ sys, double* __restrict__ posDelta) {
    const DevSys& S = *sys;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= S.numFree) return;
    const Real halfDt = (Real)(0.5 * S.dt);
    V3<Real> v = AdoptedAtoms::vel<Real>(S, a);
    V3<Real> f = AdoptedAtoms::force<Real>(S, a);
    const Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
    v = v + f * invm * halfDt;
    const V3<double> d = cvt<double>(v) * S.dt;
    double* o = posDelta + 4 * (size_t)S.ommSlot[a];
    o[0] = d.x;
    o[1] = d.y;
    o[2] = d.z;
    o[3] = 0.0;
}

template <class Real, bool EXACT, class IO>
__device__ __forceinline__ void part1_bodies(const DevSys& S, const Tile& tile) {
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    if (b >= tile.count) return;
    const int k = tile.first + b;
    const int atom0 = tile.atomFirst + b * tile.N;
    const Shape sh = body_shape(S, tile, k, atom0, tile.N);
    const int dbase = sh.row;
    Body<Real> B = load_body<Real>(S, k, sh.ks);
    V3<double> rc = load_rcm<Real>(S, k);
    drift_and_rotate<Real, EXACT>(S, B, rc);
    if (sub == 0) {
        st4<Real>(S.q, k, B.q);
        stp3<Real>(S.L, k, B.L);
        store_rcm<Real>(S, k, rc);
    }
    body_positions_any<Real, IO>(S, rotT(B.q), rc, atom0, dbase, tile.N, sub, tile.lanes);
}

// How a step kernel holds its system descriptor.  The exact-rotation kernels
// take a register copy: read through the descriptor table, its fields are
// re-read after stores that may alias the table, and each re-read waits for
// the loads in flight (heterogeneous 200k system: fused pass 22.1 -> 21.7
// us).  The NO-SQUISH kernels keep the reference: the copy's scalar
// registers spill into vector ones and cost the fused pass a wave per SIMD.
template <bool COPY> using SysView = std::conditional_t<COPY, const DevSys, const DevSys&>;

template <class Real, bool EXACT, class IO>
__global__ __launch_bounds__(EXACT ? 256 : 512) void k_part1(const DevSys* __restrict__ systems,
                                                              const Tile* __restrict__ tiles, int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        SysView<EXACT> S = systems[tile.sys];
        if (tile.kind == TILE_FREE)
            part1_free<Real, IO>(S, tile);
        else
            part1_bodies<Real, EXACT, IO>(S, tile);
    }
}

// ---------------------------------------------------------------------------
// part 2
// ---------------------------------------------------------------------------
template <class Real, class IO>
__device__ __forceinline__ void part2_free(const DevSys& S, const Tile& tile, double (&ke)[2]) {
    Real halfDt = (Real)(0.5 * S.dt);
    for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
        int a = tile.first + i;
        V3<Real> v = IO::template vel<Real>(S, a);
        V3<Real> f = IO::template force<Real>(S, a);
        Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
        v = v + f * invm * halfDt;
        // (R - saved) / dt of a constrained part 1
        if (IO::adopted && S.corrPending) v = v + ld3<Real>(S.freeCorr, S.numFree, a) * (Real)(1.0 / S.dt);
        IO::template set_vel<Real>(S, a, v);
        V3<double> vd = cvt<double>(v);
        ke[0] += dot(vd, vd) / (double)invm;
    }
}

template <class Real, class IO>
__device__ __forceinline__ void part2_bodies(const DevSys& S, const Tile& tile, double (&ke)[2]) {
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    const bool active = b < tile.count;
    const int k = tile.first + (active ? b : 0);
    const int N = tile.N;
    const int atom0 = tile.atomFirst + (active ? b : 0) * N;
    const Shape sh = body_shape(S, tile, k, atom0, N);
    const int dbase = sh.row;
    Body<Real> B = load_body<Real>(S, k, sh.ks);
    Rot<Real> R = rotT(B.q);
    V3<double> F, tau;
    body_sums_any<Real, IO>(S, R, atom0, dbase, N, sub, tile.lanes, active, F, tau);
    if (!active) return;
    const double halfDt = 0.5 * S.dt;
    V3<double> tb = axial_torque(to_body(R, tau), B.invI);
    V3<double> p = cvt<double>(B.p) + F * halfDt;
    V3<double> L = cvt<double>(B.L) + tb * halfDt;
    V3<double> w = hadamard(cvt<double>(B.invI), L);
    V3<double> vcm = p * (double)B.invM;
    V3<Real> ws = apply(R, cvt<Real>(w));
    switch (tile.lanes == 1 ? N : 0) {
    case 2: body_velocities<Real, 2, IO>(S, R, cvt<Real>(vcm), ws, atom0, dbase, N, 0, 1); break;
    case 3: body_velocities<Real, 3, IO>(S, R, cvt<Real>(vcm), ws, atom0, dbase, N, 0, 1); break;
    case 4: body_velocities<Real, 4, IO>(S, R, cvt<Real>(vcm), ws, atom0, dbase, N, 0, 1); break;
    case 6: body_velocities<Real, 6, IO>(S, R, cvt<Real>(vcm), ws, atom0, dbase, N, 0, 1); break;
    default: body_velocities<Real, 0, IO>(S, R, cvt<Real>(vcm), ws, atom0, dbase, N, sub, tile.lanes); break;
    }
    if (sub == 0) {
        ke[0] = dot(p, vcm);
        ke[1] = dot(L, w);
        // first half-kick of the next step (same F and torque)
        stp3<Real>(S.pcm, k, cvt<Real>(p + F * halfDt));
        stp3<Real>(S.L, k, cvt<Real>(L + tb * halfDt));
    }
}

template <class Real, class IO>
__global__ __launch_bounds__(512) void k_part2(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        double ke[2] = {0.0, 0.0};
        if (tile.kind == TILE_FREE)
            part2_free<Real, IO>(S, tile, ke);
        else
            part2_bodies<Real, IO>(S, tile, ke);
        block_reduce_store<2>(ke, S.kePart, tile.slot);
    }
}

// ---------------------------------------------------------------------------
// fused step boundary: part2 of step n followed by part1 of step n+1 for the
// same body, one pass over its data.  Inside RigidBodyIntegrator::step(k) the
// atom velocities and kinetic energies of the interior steps are read by
// nothing (forces depend on positions only), so this pass writes neither;
// the last step of the call ends with a plain part2.  The body state is
// rounded to storage precision at the same points as part2 + part1.
// ---------------------------------------------------------------------------
template <class Real, class IO>
__device__ __forceinline__ void fused_free(const DevSys& S, const Tile& tile) {
    Real halfDt = (Real)(0.5 * S.dt);
    for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
        int a = tile.first + i;
        V3<Real> v = IO::template vel<Real>(S, a);
        V3<Real> f = IO::template force<Real>(S, a);
        Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
        v = v + f * invm * halfDt;     // end of step n
        v = v + f * invm * halfDt;     // start of step n+1
        V3<double> r = IO::template pos<Real>(S, a);
        r = r + cvt<double>(v) * S.dt;
        IO::template set_vel<Real>(S, a, v);
        IO::template set_pos<Real>(S, a, r);
    }
}

// part2 of step n and the first half-kick of step n+1 on the body momenta
template <class Real>
__device__ __forceinline__ void fused_kicks(const DevSys& S, Body<Real>& B, const Rot<Real>& R, V3<double> F,
                                            V3<double> tau) {
    const double halfDt = 0.5 * S.dt;
    V3<double> tb = axial_torque(to_body(R, tau), B.invI);
    V3<double> p = cvt<double>(B.p) + F * halfDt;
    V3<double> L = cvt<double>(B.L) + tb * halfDt;
    B.p = cvt<Real>(p + F * halfDt);
    B.L = cvt<Real>(L + tb * halfDt);
}

template <class Real> __device__ __forceinline__ void store_body(const DevSys& S, int k, const Body<Real>& B,
                                                                 V3<double> rc) {
    st4<Real>(S.q, k, B.q);
    stp3<Real>(S.L, k, B.L);
    stp3<Real>(S.pcm, k, B.p);
    store_rcm<Real>(S, k, rc);
}

// one lane per body of NC atoms: the body-fixed coordinates stay in registers
// from the torque sum to the position rebuild, so d is read once
template <class Real, bool EXACT, int NC, class IO>
__device__ __forceinline__ void fused_body_regs(const DevSys& S, int k, int atom0, int ks, int dbase) {
    Body<Real> B = load_body<Real>(S, k, ks);
    V3<double> rc = load_rcm<Real>(S, k);
    V3<Real> dj[NC], fj[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) {
        dj[j] = ld3<Real>(S.d, S.dStride, dbase + j);
        fj[j] = IO::template force<Real>(S, atom0 + j);
    }
    Rot<Real> R = rotT(B.q);
    V3<double> F = {0.0, 0.0, 0.0}, tau = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NC; j++) {
        V3<double> dl = cvt<double>(apply(R, dj[j]));
        V3<double> f = cvt<double>(fj[j]);
        F = F + f;
        tau = tau + cross(dl, f);
    }
    fused_kicks<Real>(S, B, R, F, tau);
    drift_and_rotate<Real, EXACT>(S, B, rc);
    store_body<Real>(S, k, B, rc);
    Rot<Real> R1 = rotT(B.q);
#pragma unroll
    for (int j = 0; j < NC; j++) IO::template set_body_pos<Real>(S, atom0 + j, rc + cvt<double>(apply(R1, dj[j])));
}

template <class Real, bool EXACT, class IO>
__device__ __forceinline__ void fused_bodies(const DevSys& S, const Tile& tile) {
    const int N = tile.N;
    // (the FP64 exact rotation leaves no registers to spare: it re-reads d)
    if (!EXACT && tile.lanes == 1 && (N == 2 || N == 3 || N == 4)) {
        const int b = threadIdx.x;
        if (b >= tile.count) return;
        const int k = tile.first + b, atom0 = tile.atomFirst + b * N;
        const Shape sh = body_shape(S, tile, k, atom0, N);
        if (N == 2) fused_body_regs<Real, EXACT, 2, IO>(S, k, atom0, sh.ks, sh.row);
        else if (N == 3) fused_body_regs<Real, EXACT, 3, IO>(S, k, atom0, sh.ks, sh.row);
        else fused_body_regs<Real, EXACT, 4, IO>(S, k, atom0, sh.ks, sh.row);
        return;
    }
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    const bool active = b < tile.count;
    const int k = tile.first + (active ? b : 0);
    const int atom0 = tile.atomFirst + (active ? b : 0) * N;
    const Shape sh = body_shape(S, tile, k, atom0, N);
    const int dbase = sh.row;
    Body<Real> B = load_body<Real>(S, k, sh.ks);
    V3<double> rc = load_rcm<Real>(S, k);
    Rot<Real> R = rotT(B.q);
    V3<double> F, tau;
    body_sums_any<Real, IO>(S, R, atom0, dbase, N, sub, tile.lanes, active, F, tau);
    if (!active) return;
    fused_kicks<Real>(S, B, R, F, tau);
    drift_and_rotate<Real, EXACT>(S, B, rc);
    if (sub == 0) store_body<Real>(S, k, B, rc);
    body_positions_any<Real, IO>(S, rotT(B.q), rc, atom0, dbase, N, sub, tile.lanes);
}

template <class Real, bool EXACT, class IO>
__global__ __launch_bounds__(EXACT ? 256 : 512) void k_fused(const DevSys* __restrict__ systems,
                                                              const Tile* __restrict__ tiles, int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        SysView<EXACT> S = systems[tile.sys];
        if (tile.kind == TILE_FREE)
            fused_free<Real, IO>(S, tile);
        else
            fused_bodies<Real, EXACT, IO>(S, tile);
    }
}

// ---------------------------------------------------------------------------
// Refined energies: the CUDA platform's COMPMOD terms
// (platforms/cuda/src/kernels/rigidbodyintegrator.cu:270-292, 318-320,
// 380-383, 433-469).  The accumulators of a step depend on that step alone,
// so they are evaluated for the last step of a call only, by two kernels
// that bracket it; the step kernels themselves are the same for every
// system.  In displacement form, equal in exact arithmetic to the
// reference's 2.5 r(t) + ... and 0.5 q(t-dt) - 3 q(t) + ...:
//   rdot = dt/2 [5 v(t+dt/2) - v(t-dt/2) + 2 v(t+3dt/2)]
//   qdot = (q(t-dt) - q(t))/2 + 5 (q(t+dt) - q(t))/2 + (q(t+2dt) - q(t+dt))
// where q(t-dt) and q(t+2dt) are virtual rotations from (q, L -/+ the
// torque half-kick) over -/+dt.  Free atoms use posDot = 2 rdot of the same
// form (with the sign of v(t-dt/2) the body terms have).
//   k_refine_begin, before the step: from its starting state and the
//     current forces, a_r = 5 p(t+dt/2) - p(t-dt/2) and the two q(t) terms
//     a_q per body (q(t+dt) by the step's own kick and rotation), and
//     a_v = 5 v(t+dt/2) - v(t-dt/2) per free atom.  `folded` = the stored
//     momenta carry the step's first half-kick (after part2 or the set-up);
//     otherwise the step is a fused boundary, which applies both half-kicks.
//   k_refine_end, after the step's part2: per tile
//     [0] sum m dt/2 (a + 2 v(t+3dt/2)) . v(t+dt)   (= 6 dt KE_t refined)
//     [1] sum qdot . pi(t+dt), pi = 2 B(q) L         (= 6 dt KE_r refined)
//     (qdot's component along q, which the reference projects out, has no
//     effect: B(q)^T q = 0).
// ---------------------------------------------------------------------------
template <class Real> __device__ __forceinline__ void store_refine_body(const DevSys& S, int k, V3<double> ar,
                                                                        Q4<double> aq) {
    st4<Real>(S.refBody, 2 * k, Q4<Real>{(Real)ar.x, (Real)ar.y, (Real)ar.z, (Real)0});
    st4<Real>(S.refBody, 2 * k + 1, cvt<Real>(aq));
}

template <class Real, class IO>
__device__ __forceinline__ void refine_begin_free(const DevSys& S, const Tile& tile, bool folded) {
    Real halfDt = (Real)(0.5 * S.dt);
    for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
        int a = tile.first + i;
        V3<Real> v = IO::template vel<Real>(S, a);
        V3<Real> f = IO::template force<Real>(S, a);
        Real invm = reinterpret_cast<const Real*>(S.freeInvM)[a];
        V3<Real> dv = f * invm * halfDt;
        V3<double> av;
        if (folded) {   // v = v(t): 5 (v + dv) - (v - dv)
            av = cvt<double>(v) * 4.0 + cvt<double>(dv) * 6.0;
        } else {        // v = v(t-dt/2), v(t+dt/2) as the fused boundary forms it
            V3<Real> vp = v + dv;
            vp = vp + dv;
            av = cvt<double>(vp) * 5.0 + cvt<double>(v) * -1.0;
        }
        stp3<Real>(S.refFree, a, cvt<Real>(av));
    }
}

template <class Real, bool EXACT, class IO>
__device__ __forceinline__ void refine_begin_bodies(const DevSys& S, const Tile& tile, bool folded) {
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    const bool active = b < tile.count;
    const int k = tile.first + (active ? b : 0);
    const int atom0 = tile.atomFirst + (active ? b : 0) * tile.N;
    const Shape sh = body_shape(S, tile, k, atom0, tile.N);
    Body<Real> B = load_body<Real>(S, k, sh.ks);
    Rot<Real> R = rotT(B.q);
    V3<double> F, tau;
    body_sums_any<Real, IO>(S, R, atom0, sh.row, tile.N, sub, tile.lanes, active, F, tau);
    if (!active || sub != 0) return;
    V3<double> tb = axial_torque(to_body(R, tau), B.invI);
    V3<double> ar;
    V3<Real> Lback;
    Body<Real> B1 = B;
    if (folded) {
        // stored p = p(t+dt/2), L = L(t) + tau dt/2
        ar = cvt<double>(B.p) * 4.0 + F * S.dt;
        Lback = cvt<Real>(cvt<double>(B.L) + tb * -S.dt);
    } else {
        // stored p = p(t-dt/2), L = L(t) - tau dt/2 (after the last rotation)
        fused_kicks<Real>(S, B1, R, F, tau);
        ar = cvt<double>(B1.p) * 5.0 + cvt<double>(B.p) * -1.0;
        Lback = B.L;
    }
    Q4<Real> qm = B.q;
    free_rotation<Real, EXACT>(qm, Lback, B.invI, -S.dt, S.rotationMode);
    free_rotation<Real, EXACT>(B1.q, B1.L, B.invI, S.dt, S.rotationMode);
    Q4<double> q0 = cvt<double>(B.q) * -1.0;
    Q4<double> aq = (cvt<double>(qm) + q0) * 0.5 + (cvt<double>(B1.q) + q0) * 2.5;
    store_refine_body<Real>(S, k, ar, aq);
}

template <class Real, bool EXACT, class IO>
__global__ __launch_bounds__(EXACT ? 256 : 512) void k_refine_begin(const DevSys* __restrict__ systems,
                                                                     const Tile* __restrict__ tiles, int numTiles,
                                                                     int folded) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        if (!S.refBody) continue;
        if (tile.kind == TILE_FREE)
            refine_begin_free<Real, IO>(S, tile, folded != 0);
        else
            refine_begin_bodies<Real, EXACT, IO>(S, tile, folded != 0);
    }
}

template <class Real, class IO>
__device__ __forceinline__ void refine_end_free(const DevSys& S, const Tile& tile, double (&rk)[2]) {
    const double halfDt = 0.5 * S.dt;
    for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
        int a = tile.first + i;
        V3<double> v = cvt<double>(IO::template vel<Real>(S, a));
        V3<double> f = cvt<double>(IO::template force<Real>(S, a));
        double invm = (double)reinterpret_cast<const Real*>(S.freeInvM)[a];
        // posDot / dt = a_v + 2 v(t+3dt/2)
        V3<double> pd = cvt<double>(ldp3<Real>(S.refFree, a)) + (v + f * (invm * halfDt)) * 2.0;
        rk[0] += 0.5 * S.dt * dot(pd, v) / invm;
    }
}

template <class Real, bool EXACT, class IO>
__device__ __forceinline__ void refine_end_bodies(const DevSys& S, const Tile& tile, double (&rk)[2]) {
    const int shift = lane_shift(tile.lanes);
    const int b = threadIdx.x >> shift;
    const int sub = threadIdx.x & (tile.lanes - 1);
    const bool active = b < tile.count;
    const int k = tile.first + (active ? b : 0);
    const int atom0 = tile.atomFirst + (active ? b : 0) * tile.N;
    // stored after part2: q(t+dt), p and L with the next half-kick folded in
    const Shape sh = body_shape(S, tile, k, atom0, tile.N);
    Body<Real> B = load_body<Real>(S, k, sh.ks);
    Rot<Real> R = rotT(B.q);
    V3<double> F, tau;
    body_sums_any<Real, IO>(S, R, atom0, sh.row, tile.N, sub, tile.lanes, active, F, tau);
    if (!active || sub != 0) return;
    const double halfDt = 0.5 * S.dt;
    V3<double> tb = axial_torque(to_body(R, tau), B.invI);
    V3<double> p1 = cvt<double>(B.p) + F * -halfDt;
    V3<double> L1 = cvt<double>(B.L) + tb * -halfDt;
    Q4<Real> r0 = ld4<Real>(S.refBody, 2 * k), r1 = ld4<Real>(S.refBody, 2 * k + 1);
    // M rdot . v = dt/2 (a_r + 2 p(t+3dt/2)) . p(t+dt) / M
    V3<double> ar = {(double)r0.a, (double)r0.b, (double)r0.c};
    rk[0] = 0.5 * S.dt * (double)B.invM * dot(ar + cvt<double>(B.p) * 2.0, p1);
    // q(t+2dt): the next step's rotation from (q, L + tau dt/2)
    Q4<Real> q2 = B.q;
    V3<Real> L2 = B.L;
    free_rotation<Real, EXACT>(q2, L2, B.invI, S.dt, S.rotationMode);
    Q4<double> q1 = cvt<double>(B.q);
    Q4<double> qdot = cvt<double>(r1) + cvt<double>(q2) + q1 * -1.0;
    rk[1] = 2.0 * dot(Btmul(q1, qdot), L1);
}

template <class Real, bool EXACT, class IO>
__global__ __launch_bounds__(EXACT ? 256 : 512) void k_refine_end(const DevSys* __restrict__ systems,
                                                                   const Tile* __restrict__ tiles, int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        if (!S.refBody) continue;
        double rk[2] = {0.0, 0.0};
        if (tile.kind == TILE_FREE)
            refine_end_free<Real, IO>(S, tile, rk);
        else
            refine_end_bodies<Real, EXACT, IO>(S, tile, rk);
        block_reduce_store<2>(rk, S.refKePart, tile.slot);
    }
}

// potential-energy refinement partials (rigidbodyintegrator.cu:449-469):
// sum |F|^2/m over free atoms and |F|^2/M + tau.(I^-1 tau) over bodies, with
// the body-frame torque tau = B(q)^T C(q) tau_space of the current forces
template <class Real, class IO>
__global__ __launch_bounds__(512) void k_refine_pe(const DevSys* __restrict__ systems,
                                                    const Tile* __restrict__ tiles, int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        if (!S.refPePart) continue;
        double u[1] = {0.0};
        if (tile.kind == TILE_FREE) {
            for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
                int a = tile.first + i;
                V3<double> f = cvt<double>(IO::template force<Real>(S, a));
                u[0] += dot(f, f) * (double)reinterpret_cast<const Real*>(S.freeInvM)[a];
            }
        } else {
            const int shift = lane_shift(tile.lanes);
            const int b = threadIdx.x >> shift;
            const int sub = threadIdx.x & (tile.lanes - 1);
            const bool active = b < tile.count;
            const int k = tile.first + (active ? b : 0);
            const int atom0 = tile.atomFirst + (active ? b : 0) * tile.N;
            const Shape sh = body_shape(S, tile, k, atom0, tile.N);
            Body<Real> B = load_body<Real>(S, k, sh.ks);
            Rot<Real> R = rotT(B.q);
            V3<double> F, tau;
            body_sums_any<Real, IO>(S, R, atom0, sh.row, tile.N, sub, tile.lanes, active, F, tau);
            if (active && sub == 0) {
                V3<double> tb = axial_torque(to_body(R, tau), B.invI);
                u[0] = dot(F, F) * (double)B.invM + dot(hadamard(tb, cvt<double>(B.invI)), tb);
            }
        }
        block_reduce_store<1>(u, S.refPePart, tile.slot);
    }
}

// ---------------------------------------------------------------------------
// Distance constraints between free atoms.  The reference applies OpenMM's
// constraints to the positions after part 1 (the positions before the step
// give the bond directions) and to the velocities after part 2
// (ReferenceRigidBodyKernels.cpp:92-104); part 2's (r - saved)/dt term
// (RigidBodySystem.cpp:193-204) carries the position correction into the
// velocities.  One lane per cluster of constrained atoms: SHAKE (Ryckaert,
// Ciccotti and Berendsen 1977) and RATTLE's velocity stage (Andersen 1983),
// Gauss-Seidel over the cluster's constraints in the order given, in FP64.
// Positions are converged when every |r_ij| is within (1 +- tol) d_ij,
// velocities when every |r_ij . v_ij| <= tol |r_ij| |v_ij|.  The kicks and
// the kinetic energy are those of the unconstrained free atoms; the
// correction (r - r_drift)/dt is added right after SHAKE rather than in
// part 2 (nothing reads the velocities in between).
// ---------------------------------------------------------------------------
// Per-atom values of one cluster.  Clusters of NA <= 4 atoms (every tile
// holds clusters of one size) keep them in registers and reach atom i of a
// constraint by selects; larger ones index private memory.
template <int NA> struct ClusterSlots {
    V3<double> v[NA];
    __device__ __forceinline__ V3<double> get(int i) const {
        if constexpr (NA <= 4) {
            V3<double> o = v[0];
#pragma unroll
            for (int k = 1; k < NA; k++)
                if (i == k) o = v[k];
            return o;
        } else {
            return v[i];
        }
    }
    __device__ __forceinline__ void add(int i, V3<double> d) {
        if constexpr (NA <= 4) {
#pragma unroll
            for (int k = 0; k < NA; k++)
                if (i == k) v[k] = v[k] + d;
        } else {
            v[i] = v[i] + d;
        }
    }
};

// The constraints of one cluster: for NA <= 4 at most NA (NA - 1) / 2 (the
// distinct pairs), held in registers and read once per step; larger
// clusters read them from memory in every sweep.
template <int NA> struct ClusterCons {
    static constexpr int kMax = NA <= 4 ? NA * (NA - 1) / 2 : 1;
    // sweeps over the constraints unroll fully for the register classes; the
    // general class loops over its count
    static constexpr int kUnroll = NA <= 4 ? kMax : 1;
    int n = 0, k0 = 0, a0 = 0;
    int ci[kMax], cj[kMax];
    double cd[kMax];
    const DevSys* S = nullptr;
    __device__ __forceinline__ void load(const DevSys& sys, int first, int last, int atom0) {
        S = &sys;
        n = last - first;
        k0 = first;
        a0 = atom0;
        if constexpr (NA <= 4) {
#pragma unroll
            for (int k = 0; k < kMax; k++)
                if (k < n) {
                    ci[k] = sys.consA[first + k] - atom0;
                    cj[k] = sys.consB[first + k] - atom0;
                    cd[k] = sys.consDist[first + k];
                }
        }
    }
    // atoms and length of constraint k (0 <= k < n)
    __device__ __forceinline__ void get(int k, int& i, int& j, double& d) const {
        if constexpr (NA <= 4) {
            i = ci[k];
            j = cj[k];
            d = cd[k];
        } else {
            i = S->consA[k0 + k] - a0;
            j = S->consB[k0 + k] - a0;
            d = S->consDist[k0 + k];
        }
    }
};

template <int NA>
__device__ __forceinline__ bool shake(const DevSys& S, const ClusterCons<NA>& C, ClusterSlots<NA>& r, const ClusterSlots<NA>& ro,
                      const ClusterSlots<NA>& w) {
    const double lo = (1.0 - S.constraintTol) * (1.0 - S.constraintTol);
    const double hi = (1.0 + S.constraintTol) * (1.0 + S.constraintTol);
    for (int it = 0; it < CLUSTER_MAX_ITERATIONS; it++) {
        bool done = true;
#pragma unroll ClusterCons<NA>::kUnroll
        for (int k = 0; k < ClusterCons<NA>::kMax || NA > 4; k++) {
            if (k >= C.n) break;
            int i, j;
            double dist;
            C.get(k, i, j, dist);
            const double d2 = dist * dist;
            V3<double> s = r.get(i) - r.get(j);
            double s2 = dot(s, s);
            if (s2 >= lo * d2 && s2 <= hi * d2) continue;
            done = false;
            const double wi = w.get(i).x, wj = w.get(j).x;
            V3<double> b = ro.get(i) - ro.get(j);
            double g = (d2 - s2) / (2.0 * (wi + wj) * dot(s, b));
            r.add(i, b * (g * wi));
            r.add(j, b * (-(g * wj)));
        }
        if (done) return true;
    }
    return false;
}

template <int NA>
__device__ __forceinline__ bool rattle(const DevSys& S, const ClusterCons<NA>& C, const ClusterSlots<NA>& r, ClusterSlots<NA>& v,
                       const ClusterSlots<NA>& w) {
    const double tol2 = S.constraintTol * S.constraintTol;
    for (int it = 0; it < CLUSTER_MAX_ITERATIONS; it++) {
        bool done = true;
#pragma unroll ClusterCons<NA>::kUnroll
        for (int k = 0; k < ClusterCons<NA>::kMax || NA > 4; k++) {
            if (k >= C.n) break;
            int i, j;
            double dist;
            C.get(k, i, j, dist);
            V3<double> s = r.get(i) - r.get(j);
            V3<double> dv = v.get(i) - v.get(j);
            double rv = dot(s, dv), s2 = dot(s, s);
            if (rv * rv <= tol2 * s2 * dot(dv, dv)) continue;
            done = false;
            const double wi = w.get(i).x, wj = w.get(j).x;
            double g = -rv / ((wi + wj) * s2);
            v.add(i, s * (g * wi));
            v.add(j, s * (-(g * wj)));
        }
        if (done) return true;
    }
    return false;
}

// one cluster of n atoms (n == NA when NA <= 4, n <= NA otherwise)
template <class Real, int NA>
__device__ __forceinline__ void cluster_phase(const DevSys& S, int c, int phase, double (&ke)[2]) {
    const int a0 = S.clusterStart[c];
    const int n = NA <= 4 ? NA : S.clusterStart[c + 1] - a0;
    const Real* invm = reinterpret_cast<const Real*>(S.freeInvM);
    if (phase == CLUSTER_KE) {
        // after the set-up: as k_setup's free atoms (staged FP64 velocities)
        for (int j = 0; j < n; j++) {
            const int a = a0 + j;
            V3<double> v = S.useStageV ? V3<double>{S.stageV[3 * (size_t)a], S.stageV[3 * (size_t)a + 1],
                                                    S.stageV[3 * (size_t)a + 2]}
                                       : cvt<double>(ld3<Real>(S.vel, S.atomStride, a));
            ke[0] += dot(v, v) / (double)invm[a];
        }
        return;
    }
    const double dt = S.dt;
    const Real halfDt = (Real)(0.5 * dt);
    V3<Real> v[NA];
    ClusterSlots<NA> r, vd, w;   // w holds 1/m in x
    ClusterCons<NA> C;
    C.load(S, S.consStart[c], S.consStart[c + 1], a0);
#pragma unroll
    for (int j = 0; j < NA; j++) {
        if (NA > 4 && j >= n) break;
        v[j] = ld3<Real>(S.vel, S.atomStride, a0 + j);
        r.v[j] = load_pos<Real>(S, a0 + j);
        w.v[j] = V3<double>{1.0 / S.mass[a0 + j], 0.0, 0.0};
    }
    bool ok = true;
    if (phase == CLUSTER_PART2 || phase == CLUSTER_FUSED) {
#pragma unroll
        for (int j = 0; j < NA; j++) {
            if (NA > 4 && j >= n) break;
            V3<Real> f = ld3<Real>(S.force, S.atomStride, a0 + j);
            vd.v[j] = cvt<double>(v[j] + f * invm[a0 + j] * halfDt);
        }
        ok = rattle<NA>(S, C, r, vd, w) && ok;
#pragma unroll
        for (int j = 0; j < NA; j++) {
            if (NA > 4 && j >= n) break;
            v[j] = cvt<Real>(vd.v[j]);
            if (phase == CLUSTER_PART2) {
                V3<double> u = cvt<double>(v[j]);
                ke[0] += dot(u, u) / (double)invm[a0 + j];
            }
        }
    }
    if (phase == CLUSTER_PART1 || phase == CLUSTER_FUSED) {
        ClusterSlots<NA> ro, ru;   // before the step, after the drift
#pragma unroll
        for (int j = 0; j < NA; j++) {
            if (NA > 4 && j >= n) break;
            V3<Real> f = ld3<Real>(S.force, S.atomStride, a0 + j);
            v[j] = v[j] + f * invm[a0 + j] * halfDt;
            ro.v[j] = r.v[j];
            r.v[j] = r.v[j] + cvt<double>(v[j]) * dt;
            ru.v[j] = r.v[j];
        }
        ok = shake<NA>(S, C, r, ro, w) && ok;
        const double invDt = 1.0 / dt;
#pragma unroll
        for (int j = 0; j < NA; j++) {
            if (NA > 4 && j >= n) break;
            v[j] = cvt<Real>(cvt<double>(v[j]) + (r.v[j] - ru.v[j]) * invDt);
            store_pos<Real>(S, a0 + j, r.v[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < NA; j++) {
        if (NA > 4 && j >= n) break;
        st3<Real>(S.vel, S.atomStride, a0 + j, v[j]);
    }
    if (!ok) atomicAdd(S.constraintFail, 1u);
}

// ---- clusters of NA = 2..4 atoms: every per-atom and per-constraint value
// is a named register reached by a compile-time index (an array indexed by
// the atom numbers of a constraint would go to private memory)
template <int... I, class F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

template <class T> struct Four {
    T e0, e1, e2, e3;
    template <int J> __device__ __forceinline__ T& at() {
        if constexpr (J == 0) return e0;
        else if constexpr (J == 1) return e1;
        else if constexpr (J == 2) return e2;
        else return e3;
    }
    // element i (0 <= i < NA) as a sum of the elements weighted by exact 0/1
    // factors (selects would be turned back into an indexed load)
    template <int NA> __device__ __forceinline__ T get(int i) const {
        T o = e0 * (double)(i == 0);
        if (NA > 1) o = o + e1 * (double)(i == 1);
        if (NA > 2) o = o + e2 * (double)(i == 2);
        if (NA > 3) o = o + e3 * (double)(i == 3);
        return o;
    }
    template <int NA> __device__ __forceinline__ void add(int i, T d) {
        e0 = e0 + d * (double)(i == 0);
        if (NA > 1) e1 = e1 + d * (double)(i == 1);
        if (NA > 2) e2 = e2 + d * (double)(i == 2);
        if (NA > 3) e3 = e3 + d * (double)(i == 3);
    }
};

// the constraints of a small cluster, at most NA (NA - 1) / 2
template <int NA> struct SmallCons {
    static constexpr int kMax = NA * (NA - 1) / 2;
    int n;
    int ci[6], cj[6];
    double cd[6];
};

template <int NA>
__device__ __forceinline__ bool shake_small(const DevSys& S, const SmallCons<NA>& C, Four<V3<double>>& r,
                                            const Four<V3<double>>& ro, const Four<double>& w) {
    const double lo = (1.0 - S.constraintTol) * (1.0 - S.constraintTol);
    const double hi = (1.0 + S.constraintTol) * (1.0 + S.constraintTol);
    for (int it = 0; it < CLUSTER_MAX_ITERATIONS; it++) {
        bool done = true;
        static_for<SmallCons<NA>::kMax>([&](auto K) {
            constexpr int k = K;
            if (k >= C.n) return;
            const int i = C.ci[k], j = C.cj[k];
            const double d2 = C.cd[k] * C.cd[k];
            V3<double> s = r.template get<NA>(i) - r.template get<NA>(j);
            double s2 = dot(s, s);
            if (s2 >= lo * d2 && s2 <= hi * d2) return;
            done = false;
            const double wi = w.template get<NA>(i), wj = w.template get<NA>(j);
            V3<double> b = ro.template get<NA>(i) - ro.template get<NA>(j);
            double g = (d2 - s2) / (2.0 * (wi + wj) * dot(s, b));
            r.template add<NA>(i, b * (g * wi));
            r.template add<NA>(j, b * (-(g * wj)));
        });
        if (done) return true;
    }
    return false;
}

template <int NA>
__device__ __forceinline__ bool rattle_small(const DevSys& S, const SmallCons<NA>& C, const Four<V3<double>>& r,
                                             Four<V3<double>>& v, const Four<double>& w) {
    const double tol2 = S.constraintTol * S.constraintTol;
    for (int it = 0; it < CLUSTER_MAX_ITERATIONS; it++) {
        bool done = true;
        static_for<SmallCons<NA>::kMax>([&](auto K) {
            constexpr int k = K;
            if (k >= C.n) return;
            const int i = C.ci[k], j = C.cj[k];
            V3<double> s = r.template get<NA>(i) - r.template get<NA>(j);
            V3<double> dv = v.template get<NA>(i) - v.template get<NA>(j);
            double rv = dot(s, dv), s2 = dot(s, s);
            if (rv * rv <= tol2 * s2 * dot(dv, dv)) return;
            done = false;
            const double wi = w.template get<NA>(i), wj = w.template get<NA>(j);
            double g = -rv / ((wi + wj) * s2);
            v.template add<NA>(i, s * (g * wi));
            v.template add<NA>(j, s * (-(g * wj)));
        });
        if (done) return true;
    }
    return false;
}

// the same steps as cluster_phase, for a cluster of exactly NA atoms
template <class Real, int NA>
__device__ __forceinline__ void cluster_phase_small(const DevSys& S, int c, int phase, double (&ke)[2]) {
    const int a0 = S.clusterStart[c];
    const Real* invm = reinterpret_cast<const Real*>(S.freeInvM);
    if (phase == CLUSTER_KE) {
        static_for<NA>([&](auto J) {
            const int a = a0 + (int)J;
            V3<double> v = S.useStageV ? V3<double>{S.stageV[3 * (size_t)a], S.stageV[3 * (size_t)a + 1],
                                                    S.stageV[3 * (size_t)a + 2]}
                                       : cvt<double>(ld3<Real>(S.vel, S.atomStride, a));
            ke[0] += dot(v, v) / (double)invm[a];
        });
        return;
    }
    SmallCons<NA> C;
    const int k0 = S.consStart[c];
    C.n = S.consStart[c + 1] - k0;
    static_for<SmallCons<NA>::kMax>([&](auto K) {
        constexpr int k = K;
        if (k < C.n) {
            C.ci[k] = S.consA[k0 + k] - a0;
            C.cj[k] = S.consB[k0 + k] - a0;
            C.cd[k] = S.consDist[k0 + k];
        }
    });
    const double dt = S.dt;
    const Real halfDt = (Real)(0.5 * dt);
    Four<V3<Real>> v;
    Four<V3<double>> r, vd;
    Four<double> w;
    static_for<NA>([&](auto J) {
        constexpr int j = J;
        v.template at<j>() = ld3<Real>(S.vel, S.atomStride, a0 + j);
        r.template at<j>() = load_pos<Real>(S, a0 + j);
        w.template at<j>() = 1.0 / S.mass[a0 + j];
    });
    bool ok = true;
    if (phase == CLUSTER_PART2 || phase == CLUSTER_FUSED) {
        static_for<NA>([&](auto J) {
            constexpr int j = J;
            V3<Real> f = ld3<Real>(S.force, S.atomStride, a0 + j);
            vd.template at<j>() = cvt<double>(v.template at<j>() + f * invm[a0 + j] * halfDt);
        });
        ok = rattle_small<NA>(S, C, r, vd, w) && ok;
        static_for<NA>([&](auto J) {
            constexpr int j = J;
            v.template at<j>() = cvt<Real>(vd.template at<j>());
            if (phase == CLUSTER_PART2) {
                V3<double> u = cvt<double>(v.template at<j>());
                ke[0] += dot(u, u) / (double)invm[a0 + j];
            }
        });
    }
    if (phase == CLUSTER_PART1 || phase == CLUSTER_FUSED) {
        Four<V3<double>> ro, ru;   // before the step, after the drift
        static_for<NA>([&](auto J) {
            constexpr int j = J;
            V3<Real> f = ld3<Real>(S.force, S.atomStride, a0 + j);
            v.template at<j>() = v.template at<j>() + f * invm[a0 + j] * halfDt;
            ro.template at<j>() = r.template at<j>();
            r.template at<j>() = r.template at<j>() + cvt<double>(v.template at<j>()) * dt;
            ru.template at<j>() = r.template at<j>();
        });
        ok = shake_small<NA>(S, C, r, ro, w) && ok;
        const double invDt = 1.0 / dt;
        static_for<NA>([&](auto J) {
            constexpr int j = J;
            v.template at<j>() = cvt<Real>(cvt<double>(v.template at<j>()) + (r.template at<j>() - ru.template at<j>()) * invDt);
            store_pos<Real>(S, a0 + j, r.template at<j>());
        });
    }
    static_for<NA>([&](auto J) {
        constexpr int j = J;
        st3<Real>(S.vel, S.atomStride, a0 + j, v.template at<j>());
    });
    if (!ok) atomicAdd(S.constraintFail, 1u);
}

// tiles of clusters of NA = 2, 3 or 4 atoms (registers only), or of larger
// clusters (NA = CLUSTER_MAX_ATOMS, private memory)
template <class Real, int NA>
__global__ __launch_bounds__(256) void k_clusters(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                   int numTiles, int phase) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        double ke[3] = {0.0, 0.0, 0.0};
        if ((int)threadIdx.x < tile.count) {
            double k2[2] = {0.0, 0.0};
            if constexpr (NA <= 4) cluster_phase_small<Real, NA>(S, tile.first + threadIdx.x, phase, k2);
            else cluster_phase<Real, NA>(S, tile.first + threadIdx.x, phase, k2);
            ke[0] = k2[0];
        }
        // set-up partials carry a dof entry (the bodies'), zero here
        if (phase == CLUSTER_KE) block_reduce_store<3>(ke, S.kePart, tile.slot);
        else if (phase == CLUSTER_PART2) {
            double k2[2] = {ke[0], 0.0};
            block_reduce_store<2>(k2, S.kePart, tile.slot);
        }
    }
}

// positions as the force field reads them, as a hi part and a low part:
// body atoms (and the massless particles after them) at their fp32
// positions in mixed precision (lo = 0), free atoms at hi + lo
template <class Real, class IO = SoaAtoms>
__device__ __forceinline__ void field_pos_parts(const DevSys& S, int i, V3<Real>& hi, V3<Real>& lo) {
    lo = {(Real)0, (Real)0, (Real)0};
    if constexpr (IO::adopted) {
        const int s = S.ommSlot[i];
        Q4<Real> p = ld4<Real>(S.ommPosq, s);
        hi = {p.a, p.b, p.c};
        if (sizeof(Real) == 4 && i < S.numFree) {
            const float* c = S.ommPosqCorr + 4 * (size_t)s;
            lo = {(Real)c[0], (Real)c[1], (Real)c[2]};
        }
    } else {
        hi = ld3<Real>(S.pos, S.atomStride, i);
        if (sizeof(Real) == 4 && i < S.numFree) lo = cvt<Real>(ld3<float>(S.posLo, S.atomStride, i));
    }
}
template <class Real, class IO = SoaAtoms> __device__ __forceinline__ V3<double> force_pos(const DevSys& S, int i) {
    V3<Real> hi, lo;
    field_pos_parts<Real, IO>(S, i, hi, lo);
    return cvt<double>(hi) + cvt<double>(lo);
}
// a force into the library's array, or into the adopted fixed-point array
// (OpenMM's conversion: truncation of f * 2^32)
template <class Real, class IO> __device__ __forceinline__ void store_force(const DevSys& S, int i, V3<Real> f) {
    if constexpr (IO::adopted) {
        long long* o = S.ommForce + S.ommSlot[i];
        const size_t st = (size_t)S.ommStride;
        o[0] = (long long)((double)f.x * kFixedPointScale);
        o[st] = (long long)((double)f.y * kFixedPointScale);
        o[2 * st] = (long long)((double)f.z * kFixedPointScale);
    } else {
        st3<Real>(S.force, S.atomStride, i, f);
    }
}

// ---------------------------------------------------------------------------
// Virtual sites (OpenMM's, placed after part 1 and the constraints,
// ReferenceRigidBodyKernels.cpp:100): two- and three-particle averages and
// out-of-plane sites from their parents' positions; after the forces, each
// parent gathers the force of its sites (the transpose of d r_site /
// d r_parent) in a fixed order, and the sites' forces are cleared.  One lane
// per site or per parent.
// ---------------------------------------------------------------------------
// the local frame of a local-coordinates site: x and y are the x- and
// y-weighted sums of the parents, e_x = x/|x|, e_z = (x × y)/|x × y|,
// e_y = e_z × e_x
struct SiteFrame { V3<double> x, y, ex, ey, ez; double nx, nz; };

template <class Real>
__device__ __forceinline__ SiteFrame site_frame(const DevSys& S, const int* p, const double* w) {
    SiteFrame f;
    V3<double> r0 = force_pos<Real>(S, p[0]), r1 = force_pos<Real>(S, p[1]), r2 = force_pos<Real>(S, p[2]);
    f.x = r0 * w[3] + r1 * w[4] + r2 * w[5];
    f.y = r0 * w[6] + r1 * w[7] + r2 * w[8];
    V3<double> z = cross(f.x, f.y);
    f.nx = sqrt(dot(f.x, f.x));
    f.nz = sqrt(dot(z, z));
    f.ex = f.x * (1.0 / f.nx);
    f.ez = z * (1.0 / f.nz);
    f.ey = cross(f.ez, f.ex);
    return f;
}

template <class Real>
__global__ __launch_bounds__(256) void k_site_positions(const DevSys* __restrict__ sys) {
    const DevSys& S = *sys;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S.numSites) return;
    const int* p = S.siteParent + 3 * (size_t)s;
    const double* w = S.siteW + 12 * (size_t)s;
    V3<double> r1 = force_pos<Real>(S, p[0]), r2 = force_pos<Real>(S, p[1]), x;
    const int type = S.siteType[s];
    if (type == SITE_LOCAL_COORDINATES) {
        SiteFrame f = site_frame<Real>(S, p, w);
        V3<double> o = r1 * w[0] + r2 * w[1] + force_pos<Real>(S, p[2]) * w[2];
        x = o + f.ex * w[9] + f.ey * w[10] + f.ez * w[11];
    } else if (type == SITE_TWO_AVERAGE) {
        x = r1 * w[0] + r2 * w[1];
    } else if (type == SITE_THREE_AVERAGE) {
        x = r1 * w[0] + r2 * w[1] + force_pos<Real>(S, p[2]) * w[2];
    } else {
        V3<double> r12 = r2 - r1, r13 = force_pos<Real>(S, p[2]) - r1;
        x = r1 + r12 * w[0] + r13 * w[1] + cross(r12, r13) * w[2];
    }
    store_pos<Real>(S, S.siteAtom[s], x);
}

template <class Real>
__global__ __launch_bounds__(256) void k_site_forces(const DevSys* __restrict__ sys) {
    const DevSys& S = *sys;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= S.numSiteParents) return;
    const int a = S.parentAtom[g];
    V3<double> f = cvt<double>(ld3<Real>(S.force, S.atomStride, a));
    for (int e = S.parentStart[g]; e < S.parentStart[g + 1]; e++) {
        const int s = S.parentSite[e], role = S.parentRole[e];
        const int* p = S.siteParent + 3 * (size_t)s;
        const double* w = S.siteW + 12 * (size_t)s;
        V3<double> fs = cvt<double>(ld3<Real>(S.force, S.atomStride, S.siteAtom[s]));
        const int type = S.siteType[s];
        if (type == SITE_TWO_AVERAGE || type == SITE_THREE_AVERAGE) {
            f = f + fs * w[role];
        } else if (type == SITE_LOCAL_COORDINATES) {
            // ow F + xw Gx + yw Gy, Gx and Gy the transposed derivatives of
            // the site with respect to x and y
            SiteFrame fr = site_frame<Real>(S, p, w);
            auto Jx = [&](V3<double> a) { return (a - fr.ex * dot(fr.ex, a)) * (1.0 / fr.nx); };
            auto Jz = [&](V3<double> a) { return (a - fr.ez * dot(fr.ez, a)) * (1.0 / fr.nz); };
            V3<double> jzexf = Jz(cross(fr.ex, fs)), jzf = Jz(fs);
            V3<double> gx = Jx(fs) * w[9] + (cross(fr.y, jzexf) - Jx(cross(fr.ez, fs))) * w[10] + cross(fr.y, jzf) * w[11];
            V3<double> gy = cross(fr.x, jzexf) * (-w[10]) - cross(fr.x, jzf) * w[11];
            f = f + (fs * w[role] + gx * w[3 + role] + gy * w[6 + role]);
        } else {
            V3<double> r1 = force_pos<Real>(S, p[0]);
            V3<double> r12 = force_pos<Real>(S, p[1]) - r1, r13 = force_pos<Real>(S, p[2]) - r1;
            V3<double> f2 = fs * w[0] + cross(r13, fs) * w[2];
            V3<double> f3 = fs * w[1] + cross(fs, r12) * w[2];
            f = f + (role == 0 ? fs - f2 - f3 : role == 1 ? f2 : f3);
        }
    }
    st3<Real>(S.force, S.atomStride, a, cvt<Real>(f));
}

template <class Real>
__global__ __launch_bounds__(256) void k_site_clear(const DevSys* __restrict__ sys) {
    const DevSys& S = *sys;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S.numSites) return;
    st3<Real>(S.force, S.atomStride, S.siteAtom[s], V3<Real>{(Real)0, (Real)0, (Real)0});
}

// ---------------------------------------------------------------------------
// force field: harmonic tether + harmonic bonds, one lane per atom.  Body
// atoms are read at their fp32 positions (as a mixed-precision force field
// reads fp32 coordinates), free atoms at hi + lo; the tether's anchor r0 is
// held at the state precision (fp32 in mixed precision, as a mixed-precision
// force field holds its parameters), so an atom moves 36 bytes: its
// position, r0 and its force.  With ENERGY the tiles' potential-energy
// partials too; a step asks for forces only (calcForcesAndEnergy(true,
// false), ReferenceRigidBodyKernels.cpp:101), so a sequence of steps takes
// the energy at its last evaluation alone.
// ---------------------------------------------------------------------------

template <class Real, class IO, bool ENERGY>
__global__ __launch_bounds__(512) void k_forces(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                 int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        // (a copy: read through the table, the descriptor's fields would be
        // re-read after every store, which may alias it, and each re-read
        // waits for the atom loads in flight)
        const DevSys S = systems[tile.sys];
        double e[1] = {0.0};
        for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
            int a = tile.first + i;
            V3<Real> f = {(Real)0, (Real)0, (Real)0};
            if (S.tetherScale != 0.0) {
                Real k = S.tetherUniform ? (Real)S.tetherK0 : reinterpret_cast<const Real*>(S.tetherK)[a];
                V3<Real> ph, pl;
                field_pos_parts<Real, IO>(S, a, ph, pl);
                V3<Real> dx = (ph - ld3<Real>(S.r0, S.atomStride, a)) + pl;
                f = dx * (-k);
                if (ENERGY) {
                    V3<double> dd = cvt<double>(dx);
                    e[0] += 0.5 * (double)k * dot(dd, dd);
                }
            }
            if (S.bondStart) {
                int e0 = S.bondStart[a], e1 = S.bondStart[a + 1];
                if (e0 < e1) {
                    V3<double> pa = force_pos<Real, IO>(S, a);
                    V3<double> fd = cvt<double>(f);
                    for (int b = e0; b < e1; b++) {
                        V3<double> dr = force_pos<Real, IO>(S, S.bondOther[b]) - pa;
                        double r = sqrt(dot(dr, dr));
                        double ext = r - S.bondLen[b];
                        fd = fd + dr * (S.bondK[b] * ext / r);
                        if (ENERGY) e[0] += 0.25 * S.bondK[b] * ext * ext;
                    }
                    f = cvt<Real>(fd);
                }
            }
            store_force<Real, IO>(S, a, f);
        }
        if (ENERGY) block_reduce_store<1>(e, S.pePart, tile.slot);
    }
}

// deterministic sum of `nslots` 4-double partials of each system; one block per system
__global__ __launch_bounds__(256) void k_reduce(const DevSys* __restrict__ systems, const int* __restrict__ nslots,
                                                 int which, double* __restrict__ out) {
    const DevSys& S = systems[blockIdx.x];
    const double* part = which == 0 ? S.kePart : which == 1 ? S.pePart : which == 2 ? S.refKePart : S.refPePart;
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

// Timing of the step kernels: with events set by set_step_kernel_events
// (per thread), the next step kernel is launched with them attached to its
// dispatch (hipExtLaunchKernelGGL), so they record its own start and end;
// both are then cleared.  Without them, a plain launch.
namespace {
thread_local hipEvent_t t_kernelStart = nullptr, t_kernelStop = nullptr;

template <class... KArgs, class... Args>
void step_kernel(void (*kernel)(KArgs...), int grid, int block, hipStream_t stream, Args... args) {
    if (t_kernelStart) {
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, t_kernelStart, t_kernelStop, 0, args...);
        t_kernelStart = t_kernelStop = nullptr;
    } else {
        kernel<<<grid, block, 0, stream>>>(args...);
    }
}
}  // namespace

void set_step_kernel_events(hipEvent_t start, hipEvent_t stop) {
    t_kernelStart = start;
    t_kernelStop = stop;
}

bool clear_step_kernel_events() {
    const bool pending = t_kernelStart != nullptr;
    t_kernelStart = t_kernelStop = nullptr;
    return pending;
}

// the step kernels for the atom layout (library arrays or adopted arrays),
// state precision and rotation kind
template <class IO>
static void part1_for(int precision, int rotationMode, int grid, int block, hipStream_t stream,
                      const DevSys* systems, const Tile* tiles, int numTiles) {
    if (precision == 0) {
        if (rotationMode == 0) step_kernel(k_part1<float, true, IO>, grid, block, stream, systems, tiles, numTiles);
        else step_kernel(k_part1<float, false, IO>, grid, block, stream, systems, tiles, numTiles);
    } else {
        if (rotationMode == 0) step_kernel(k_part1<double, true, IO>, grid, block, stream, systems, tiles, numTiles);
        else step_kernel(k_part1<double, false, IO>, grid, block, stream, systems, tiles, numTiles);
    }
}
template <class IO>
static void fused_for(int precision, int rotationMode, int grid, int block, hipStream_t stream,
                      const DevSys* systems, const Tile* tiles, int numTiles) {
    if (precision == 0) {
        if (rotationMode == 0) step_kernel(k_fused<float, true, IO>, grid, block, stream, systems, tiles, numTiles);
        else step_kernel(k_fused<float, false, IO>, grid, block, stream, systems, tiles, numTiles);
    } else {
        if (rotationMode == 0) step_kernel(k_fused<double, true, IO>, grid, block, stream, systems, tiles, numTiles);
        else step_kernel(k_fused<double, false, IO>, grid, block, stream, systems, tiles, numTiles);
    }
}

hipError_t launch_part1(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                        int block, int maxBlocks, hipStream_t stream, int adopted) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (adopted) part1_for<AdoptedAtoms>(precision, rotationMode, grid, block, stream, systems, tiles, numTiles);
    else part1_for<SoaAtoms>(precision, rotationMode, grid, block, stream, systems, tiles, numTiles);
    return hipGetLastError();
}

hipError_t launch_fused(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                        int block, int maxBlocks, hipStream_t stream, int adopted) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (adopted) fused_for<AdoptedAtoms>(precision, rotationMode, grid, block, stream, systems, tiles, numTiles);
    else fused_for<SoaAtoms>(precision, rotationMode, grid, block, stream, systems, tiles, numTiles);
    return hipGetLastError();
}

// The free-atom tiles' kinetic-energy partials again, from the adopted
// velocities as they are now (after the caller's velocity constraints), into
// the slots part 2 filled; body tiles keep theirs.
template <class Real>
__global__ __launch_bounds__(512) void k_free_ke(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                  int numTiles) {
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        if (tile.kind != TILE_FREE) continue;
        const DevSys& S = systems[tile.sys];
        double ke[2] = {0.0, 0.0};
        for (int i = threadIdx.x; i < tile.count; i += blockDim.x) {
            const int a = tile.first + i;
            V3<double> vd = cvt<double>(AdoptedAtoms::vel<Real>(S, a));
            ke[0] += dot(vd, vd) / (double)reinterpret_cast<const Real*>(S.freeInvM)[a];
        }
        block_reduce_store<2>(ke, S.kePart, tile.slot);
    }
}

hipError_t launch_free_ke(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                          hipStream_t stream) {
    if (numTiles <= 0) return hipSuccess;
    if (precision == 0) k_free_ke<float><<<numTiles, block, 0, stream>>>(systems, tiles, numTiles);
    else k_free_ke<double><<<numTiles, block, 0, stream>>>(systems, tiles, numTiles);
    return hipGetLastError();
}

hipError_t launch_free_delta(int precision, const DevSys* sys, int numFree, double* posDelta, hipStream_t stream) {
    if (numFree <= 0) return hipSuccess;
    const int grid = (numFree + 255) / 256;
    if (precision == 0) k_free_delta<float><<<grid, 256, 0, stream>>>(sys, posDelta);
    else k_free_delta<double><<<grid, 256, 0, stream>>>(sys, posDelta);
    return hipGetLastError();
}

hipError_t launch_part2(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                        int maxBlocks, hipStream_t stream, int adopted) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) {
        if (adopted) step_kernel(k_part2<float, AdoptedAtoms>, grid, block, stream, systems, tiles, numTiles);
        else step_kernel(k_part2<float, SoaAtoms>, grid, block, stream, systems, tiles, numTiles);
    } else {
        if (adopted) step_kernel(k_part2<double, AdoptedAtoms>, grid, block, stream, systems, tiles, numTiles);
        else step_kernel(k_part2<double, SoaAtoms>, grid, block, stream, systems, tiles, numTiles);
    }
    return hipGetLastError();
}

template <class Real>
static void launch_cluster_classes(const DevSys* systems, const Tile* tiles, const int* counts, int phase, int block,
                                   hipStream_t stream) {
    if (counts[0]) k_clusters<Real, 2><<<counts[0], block, 0, stream>>>(systems, tiles, counts[0], phase);
    tiles += counts[0];
    if (counts[1]) k_clusters<Real, 3><<<counts[1], block, 0, stream>>>(systems, tiles, counts[1], phase);
    tiles += counts[1];
    if (counts[2]) k_clusters<Real, 4><<<counts[2], block, 0, stream>>>(systems, tiles, counts[2], phase);
    tiles += counts[2];
    if (counts[3])
        k_clusters<Real, CLUSTER_MAX_ATOMS><<<counts[3], block, 0, stream>>>(systems, tiles, counts[3], phase);
}

hipError_t launch_clusters(int precision, const DevSys* systems, const Tile* tiles, const int* counts, int phase,
                           int block, hipStream_t stream) {
    if (block > 256) return hipErrorInvalidValue;
    for (int i = 0; i < 4; i++)
        if (counts[i] < 0) return hipErrorInvalidValue;
    if (precision == 0) launch_cluster_classes<float>(systems, tiles, counts, phase, block, stream);
    else launch_cluster_classes<double>(systems, tiles, counts, phase, block, stream);
    return hipGetLastError();
}

hipError_t launch_site_positions(int precision, const DevSys* sys, int numSites, hipStream_t stream) {
    if (numSites <= 0) return hipSuccess;
    const int grid = (numSites + 255) / 256;
    if (precision == 0) k_site_positions<float><<<grid, 256, 0, stream>>>(sys);
    else k_site_positions<double><<<grid, 256, 0, stream>>>(sys);
    return hipGetLastError();
}

hipError_t launch_site_forces(int precision, const DevSys* sys, int numSites, int numParents, hipStream_t stream) {
    if (numSites <= 0) return hipSuccess;
    const int gp = (numParents + 255) / 256, gs = (numSites + 255) / 256;
    if (precision == 0) {
        if (gp) k_site_forces<float><<<gp, 256, 0, stream>>>(sys);
        k_site_clear<float><<<gs, 256, 0, stream>>>(sys);
    } else {
        if (gp) k_site_forces<double><<<gp, 256, 0, stream>>>(sys);
        k_site_clear<double><<<gs, 256, 0, stream>>>(sys);
    }
    return hipGetLastError();
}

template <class Real, bool EXACT>
static void refine_begin_for(int adopted, int grid, int block, hipStream_t stream, const DevSys* systems,
                             const Tile* tiles, int numTiles, int folded) {
    if (adopted) k_refine_begin<Real, EXACT, AdoptedAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles, folded);
    else k_refine_begin<Real, EXACT, SoaAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles, folded);
}

hipError_t launch_refine_begin(int precision, int rotationMode, const DevSys* systems, const Tile* tiles,
                               int numTiles, int folded, int block, int maxBlocks, hipStream_t stream, int adopted) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) {
        if (rotationMode == 0) refine_begin_for<float, true>(adopted, grid, block, stream, systems, tiles, numTiles, folded);
        else refine_begin_for<float, false>(adopted, grid, block, stream, systems, tiles, numTiles, folded);
    } else {
        if (rotationMode == 0) refine_begin_for<double, true>(adopted, grid, block, stream, systems, tiles, numTiles, folded);
        else refine_begin_for<double, false>(adopted, grid, block, stream, systems, tiles, numTiles, folded);
    }
    return hipGetLastError();
}

template <class Real, bool EXACT>
static void refine_end_for(int adopted, int grid, int block, hipStream_t stream, const DevSys* systems,
                           const Tile* tiles, int numTiles) {
    if (adopted) k_refine_end<Real, EXACT, AdoptedAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    else k_refine_end<Real, EXACT, SoaAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
}

hipError_t launch_refine_end(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                             int block, int maxBlocks, hipStream_t stream, int adopted) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) {
        if (rotationMode == 0) refine_end_for<float, true>(adopted, grid, block, stream, systems, tiles, numTiles);
        else refine_end_for<float, false>(adopted, grid, block, stream, systems, tiles, numTiles);
    } else {
        if (rotationMode == 0) refine_end_for<double, true>(adopted, grid, block, stream, systems, tiles, numTiles);
        else refine_end_for<double, false>(adopted, grid, block, stream, systems, tiles, numTiles);
    }
    return hipGetLastError();
}

hipError_t launch_refine_pe(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                            int maxBlocks, hipStream_t stream, int adopted) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) {
        if (adopted) k_refine_pe<float, AdoptedAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
        else k_refine_pe<float, SoaAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    } else {
        if (adopted) k_refine_pe<double, AdoptedAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
        else k_refine_pe<double, SoaAtoms><<<grid, block, 0, stream>>>(systems, tiles, numTiles);
    }
    return hipGetLastError();
}

template <class Real, class IO>
static void forces_for(bool energy, int grid, int block, hipStream_t stream, const DevSys* systems, const Tile* tiles,
                       int numTiles) {
    if (energy) step_kernel(k_forces<Real, IO, true>, grid, block, stream, systems, tiles, numTiles);
    else step_kernel(k_forces<Real, IO, false>, grid, block, stream, systems, tiles, numTiles);
}

hipError_t launch_forces(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                         int maxBlocks, hipStream_t stream, int adopted, int energy) {
    int grid = grid_for(numTiles, maxBlocks);
    if (grid == 0) return hipSuccess;
    if (precision == 0) {
        if (adopted) forces_for<float, AdoptedAtoms>(energy, grid, block, stream, systems, tiles, numTiles);
        else forces_for<float, SoaAtoms>(energy, grid, block, stream, systems, tiles, numTiles);
    } else {
        if (adopted) forces_for<double, AdoptedAtoms>(energy, grid, block, stream, systems, tiles, numTiles);
        else forces_for<double, SoaAtoms>(energy, grid, block, stream, systems, tiles, numTiles);
    }
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
