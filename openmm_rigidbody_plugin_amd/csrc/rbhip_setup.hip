// Body-state build and state conversion kernels.  This translation unit is
// compiled with -ffp-contract=off: the body-state build runs in FP64 with the
// operation order of the reference's host code, so the device starts from the
// same body frames, quaternions and momenta the FP64 CPU path computes.
//
//   k_setup        RigidBody::buildGeometry (openmmapi/src/RigidBody.cpp:65-116)
//                  and RigidBody::buildDynamics (RigidBody.cpp:123-142) for
//                  each body; KE and dof partials; first half-kick of the
//                  next step (RigidBodySystem.cpp:177-179)
//   k_body_states  RigidBody fields for download (half-kick removed)
//   k_rekick       re-applies the folded half-kick after a step-size change
//   k_stage_in/out FP64 host-order staging <-> device precision state
//   k_body_atoms_out  FP64 body-atom positions r_cm + A(q)^T d_j rebuilt from
//                  the body state (the device keeps them in fp32 only)
//   k_match_frames bodies of one shape: shared body-fixed coordinates
#include <hip/hip_runtime.h>

#include "rbhip_access.h"
#include "rbhip_kernels.h"

namespace rbhip {

// ---------------------------------------------------------------------------
// geometry helpers (reference RigidBody.cpp:26-58)
// ---------------------------------------------------------------------------
// unit vector orthogonal to u: Projection(u) e_m / |.|, m = index of the
// smallest |u_k| (the reference takes the smallest signed component, which
// is parallel to u when u = -e_k and gives 0/0; with |u_m| <= 1/sqrt(3) the
// projection keeps a length >= sqrt(2/3))
__device__ inline V3<double> orthonormal(V3<double> u) {
    int m = fabs(u.x) < fabs(u.y) ? 0 : 1;
    double um = m == 0 ? u.x : u.y;
    if (fabs(u.z) < fabs(um)) { m = 2; um = u.z; }
    double xtx = dot(u, u);
    V3<double> v = {(-u.x) * um, (-u.y) * um, (-u.z) * um};
    if (m == 0) v.x += xtx;
    else if (m == 1) v.y += xtx;
    else v.z += xtx;
    return v * (1.0 / sqrt(dot(v, v)));
}

// inertia += Projection(delta) * m   (MatVec.cpp:555-564)
__device__ inline void add_projection(M3& I, V3<double> x, double m) {
    double xtx = dot(x, x);
    V3<double> nx = {-x.x, -x.y, -x.z};
    V3<double> r0 = nx * x.x, r1 = nx * x.y, r2 = nx * x.z;
    r0.x += xtx; r1.y += xtx; r2.z += xtx;
    I.r[0] = I.r[0] + r0 * m;
    I.r[1] = I.r[1] + r1 * m;
    I.r[2] = I.r[2] + r2 * m;
}

__device__ inline V3<double> m3_apply(const M3& A, V3<double> x) {
    return {dot(A.r[0], x), dot(A.r[1], x), dot(A.r[2], x)};
}

// body-frame torque A(q) tau without the axial part of a linear body
__device__ inline V3<double> body_torque(Q4<double> q, V3<double> tau, double invIz) {
    V3<double> tb = Amul(q, tau);
    if (invIz == 0.0) tb.z = 0.0;
    return tb;
}

template <class Real> __device__ __forceinline__ V3<double> setup_pos(const DevSys& S, int i) {
    if (S.useStageR) return V3<double>{S.stageR[3 * (size_t)i], S.stageR[3 * (size_t)i + 1], S.stageR[3 * (size_t)i + 2]};
    return load_pos<Real>(S, i);
}
template <class Real> __device__ __forceinline__ V3<double> setup_vel(const DevSys& S, int i) {
    if (S.useStageV) return V3<double>{S.stageV[3 * (size_t)i], S.stageV[3 * (size_t)i + 1], S.stageV[3 * (size_t)i + 2]};
    return cvt<double>(ld3<Real>(S.vel, S.atomStride, i));
}

template <class Real>
__device__ void setup_body(const DevSys& S, int k, int N, int atom0, int geometry, double (&red)[3]) {
    const int dbase = atom0 - S.numFree;
    const double TOL = 1.0e-5;
    double mass = 0.0;
    for (int j = 0; j < N; j++) mass += S.mass[atom0 + j];
    const double invMass = 1.0 / mass;
    Q4<double> q;
    V3<double> invI, rcm = {0.0, 0.0, 0.0}, F = {0.0, 0.0, 0.0}, tau = {0.0, 0.0, 0.0};
    M3 A;
    if (geometry) {
        for (int j = 0; j < N; j++) rcm = rcm + setup_pos<Real>(S, atom0 + j) * S.mass[atom0 + j];
        rcm = rcm * (1.0 / mass);
        // collinearity (RigidBody.cpp:26-45): the reference keeps the last atom
        // farther than atom 0 as the direction candidate
        V3<double> dl0 = setup_pos<Real>(S, atom0) - rcm;
        double d0 = dot(dl0, dl0), d2max = d0;
        int jmax = 0;
        for (int j = 1; j < N; j++) {
            V3<double> dl = setup_pos<Real>(S, atom0 + j) - rcm;
            double d2 = dot(dl, dl);
            if (d2 > d0) { jmax = j; d2max = d2; }
        }
        V3<double> u = (setup_pos<Real>(S, atom0 + jmax) - rcm) * (1.0 / sqrt(d2max));
        bool lin = true;
        for (int j = 0; lin && j < N; j++) {
            V3<double> dl = setup_pos<Real>(S, atom0 + j) - rcm;
            double d2 = dot(dl, dl);
            double ud = dot(u, dl);
            lin = d2 < TOL * d2max || fabs(ud * ud / d2 - 1.0) < TOL;
        }
        if (lin) {
            double moi = 0.0;
            for (int j = 0; j < N; j++) {
                V3<double> dl = setup_pos<Real>(S, atom0 + j) - rcm;
                moi += S.mass[atom0 + j] * dot(dl, dl);
            }
            invI = {1.0 / moi, 1.0 / moi, 0.0};
            V3<double> v = orthonormal(u);
            // frame rows (v, u x v, u): the molecular axis is body axis 3, the
            // zero principal moment (the reference stores the transpose,
            // RigidBody.cpp:96, which breaks I = (MoI, MoI, 0))
            A.r[0] = v;
            A.r[1] = cross(u, v);
            A.r[2] = u;
        } else {
            M3 inertia;
            inertia.r[0] = inertia.r[1] = inertia.r[2] = V3<double>{0.0, 0.0, 0.0};
            for (int j = 0; j < N; j++)
                add_projection(inertia, setup_pos<Real>(S, atom0 + j) - rcm, S.mass[atom0 + j]);
            V3<double> I = eigenvalues3(inertia);
            invI = {1.0 / I.x, 1.0 / I.y, 1.0 / I.z};
            A = eigenvectors3(inertia, I);
        }
        q = quat_from_matrix(A);
        S.dOff[k] = k;   // its own rows and inertia until k_match_frames finds a shared shape
        for (int j = 0; j < N; j++) {
            V3<double> dl = setup_pos<Real>(S, atom0 + j) - rcm;
            st3<Real>(S.d, S.dStride, dbase + j, cvt<Real>(m3_apply(A, dl)));
            V3<double> f = cvt<double>(ld3<Real>(S.force, S.atomStride, atom0 + j));
            F = F + f;
            tau = tau + cross(dl, f);
        }
        const int st = S.bodyStride;
        S.rcm[k] = rcm.x;
        S.rcm[k + st] = rcm.y;
        S.rcm[k + 2 * st] = rcm.z;
        st4<Real>(S.iI, k, Q4<Real>{(Real)invI.x, (Real)invI.y, (Real)invI.z, (Real)invMass});
        st4<Real>(S.q, k, cvt<Real>(q));
    } else {
        q = cvt<double>(ld4<Real>(S.q, k));
        Q4<Real> iI = ld4<Real>(S.iI, k);
        invI = {(double)iI.a, (double)iI.b, (double)iI.c};
        for (int j = 0; j < N; j++) {
            V3<double> dl = Atmul(q, cvt<double>(ld3<Real>(S.d, S.dStride, dbase + j)));
            V3<double> f = cvt<double>(ld3<Real>(S.force, S.atomStride, atom0 + j));
            F = F + f;
            tau = tau + cross(dl, f);
        }
    }
    const double dof = invI.z == 0.0 ? 5.0 : 6.0;
    // dynamics (RigidBody.cpp:123-142), centre-of-mass momentum rebuilt from zero
    V3<double> pcm = {0.0, 0.0, 0.0};
    for (int j = 0; j < N; j++) pcm = pcm + setup_vel<Real>(S, atom0 + j) * S.mass[atom0 + j];
    V3<double> vcm = pcm * (1.0 / mass);
    double twoKt = dot(pcm, vcm);
    V3<double> L = {0.0, 0.0, 0.0};
    for (int j = 0; j < N; j++) {
        V3<double> dj;
        if (geometry) dj = m3_apply(A, setup_pos<Real>(S, atom0 + j) - rcm);
        else dj = cvt<double>(ld3<Real>(S.d, S.dStride, dbase + j));
        L = L + cross(dj, Amul(q, setup_vel<Real>(S, atom0 + j) - vcm) * S.mass[atom0 + j]);
    }
    // a linear body has no angular momentum about its axis (body axis 3)
    if (invI.z == 0.0) L.z = 0.0;
    double twoKr = dot(L, hadamard(invI, L));
    // first half-kick of the next step (RigidBodySystem.cpp:177-178); the
    // reference's pi += C(q) tau dt is L += A(q) tau dt/2 for pi = 2 B(q) L
    const double halfDt = 0.5 * S.dt;
    pcm = pcm + F * halfDt;
    L = L + body_torque(q, tau, invI.z) * halfDt;
    stp3<Real>(S.pcm, k, cvt<Real>(pcm));
    stp3<Real>(S.L, k, cvt<Real>(L));
    red[0] = twoKt;
    red[1] = twoKr;
    red[2] = dof;
}

__device__ inline double wave_sum_s(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <class Real>
__global__ __launch_bounds__(256) void k_setup(const DevSys* __restrict__ systems, const Tile* __restrict__ tiles,
                                                 int numTiles, int geometry) {
    __shared__ double red[4][4];
    for (int t = blockIdx.x; t < numTiles; t += gridDim.x) {
        const Tile tile = tiles[t];
        const DevSys& S = systems[tile.sys];
        double v[3] = {0.0, 0.0, 0.0};
        int i = threadIdx.x;
        if (tile.kind == TILE_FREE) {
            for (int f = i; f < tile.count; f += blockDim.x) {
                int a = tile.first + f;
                V3<double> vel = setup_vel<Real>(S, a);
                v[0] += dot(vel, vel) / (double)reinterpret_cast<const Real*>(S.freeInvM)[a];
            }
        } else if (i < tile.count) {
            setup_body<Real>(S, tile.first + i, tile.N, tile.atomFirst + i * tile.N, geometry, v);
        }
        int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int c = 0; c < 3; c++) {
            double s = wave_sum_s(v[c]);
            if (lane == 0) red[wave][c] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int nw = (blockDim.x + 63) >> 6;
            for (int c = 0; c < 3; c++) {
                double s = 0.0;
                for (int w = 0; w < nw; w++) s += red[w][c];
                S.kePart[4 * (size_t)tile.slot + c] = s;
            }
            S.kePart[4 * (size_t)tile.slot + 3] = 0.0;
        }
        __syncthreads();
    }
}

// FP64 staging -> device state (the library's arrays or the adopted ones):
// what = 0 positions, 1 velocities (rounded to the state precision), 2
// forces (library arrays only).  Internal atom i reads row orig[i] of `in`
// (row i when orig is null) and, with `keep`, leaves its FP64 row there.
template <class Real, class IO>
__global__ void k_stage_in(const DevSys* __restrict__ sys, const double* __restrict__ in, int n, int what,
                           const int* __restrict__ orig, double* __restrict__ keep) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t r = orig ? (size_t)orig[i] : (size_t)i;
    V3<double> x = {in[3 * r], in[3 * r + 1], in[3 * r + 2]};
    if (keep) {
        keep[3 * (size_t)i] = x.x;
        keep[3 * (size_t)i + 1] = x.y;
        keep[3 * (size_t)i + 2] = x.z;
    }
    if (what == 0) IO::template set_pos<Real>(*sys, i, x);
    else if (what == 1) IO::template set_vel<Real>(*sys, i, cvt<Real>(x));
    else st3<Real>(sys->force, sys->atomStride, i, cvt<Real>(x));
}

// device state -> FP64 staging: what = 0 positions, 1 velocities, 2 forces
// (adopted forces at their fixed-point value); internal atom i goes to row
// orig[i] of `out` (row i when orig is null)
template <class Real, class IO>
__global__ void k_stage_out(const DevSys* __restrict__ sys, double* __restrict__ out, int n, int what,
                            const int* __restrict__ orig) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevSys& S = *sys;
    V3<double> x;
    if (what == 0) {
        x = IO::template pos<Real>(S, i);
    } else if (what == 1) {
        if constexpr (IO::adopted) {   // velm holds FP64 values
            const double* v = S.ommVelm + 4 * (size_t)S.ommSlot[i];
            x = {v[0], v[1], v[2]};
        } else {
            x = cvt<double>(IO::template vel<Real>(S, i));
        }
    } else if constexpr (IO::adopted) {
        const long long* f = S.ommForce + S.ommSlot[i];
        const size_t st = (size_t)S.ommStride;
        const double u = 1.0 / kFixedPointScale;
        x = {(double)f[0] * u, (double)f[st] * u, (double)f[2 * st] * u};
    } else {
        x = cvt<double>(ld3<Real>(S.force, S.atomStride, i));
    }
    const size_t r = orig ? (size_t)orig[i] : (size_t)i;
    out[3 * r] = x.x;
    out[3 * r + 1] = x.y;
    out[3 * r + 2] = x.z;
}

// tether anchors (what = 0: FP64 xyz rows) or spring constants (what = 1:
// one FP64 value per atom), row orig[i] for internal atom i, at the state
// precision
template <class Real>
__global__ void k_tether_in(const DevSys* __restrict__ sys, const double* __restrict__ in, int n, int what,
                            const int* __restrict__ orig) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevSys& S = *sys;
    const size_t r = (size_t)orig[i];
    if (what == 0) {
        Real* r0 = static_cast<Real*>(S.r0);
        const size_t as = (size_t)S.atomStride;
        r0[i] = (Real)in[3 * r];
        r0[i + as] = (Real)in[3 * r + 1];
        r0[i + 2 * as] = (Real)in[3 * r + 2];
    } else {
        static_cast<Real*>(S.tetherK)[i] = (Real)in[r];
    }
}

// adopted fixed-point forces -> the library's force array at the state
// precision (the body-state build, the re-kick and the body-state download
// read it there)
template <class Real>
__global__ void k_adopted_forces_in(const DevSys* __restrict__ sys, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    st3<Real>(sys->force, sys->atomStride, i, AdoptedAtoms::force<Real>(*sys, i));
}

// moves the adopted force array's entries of the n internal atoms from
// their old slots to new ones (the atom order of the arrays changed): a
// gather into `tmp` (3 per atom) with the old map, then a scatter with the
// new one
__global__ void k_adopted_gather_forces(const DevSys* __restrict__ sys, long long* __restrict__ tmp, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevSys& S = *sys;
    const long long* f = S.ommForce + S.ommSlot[i];
    const size_t st = (size_t)S.ommStride;
    tmp[3 * (size_t)i] = f[0];
    tmp[3 * (size_t)i + 1] = f[st];
    tmp[3 * (size_t)i + 2] = f[2 * st];
}
__global__ void k_adopted_scatter_forces(const DevSys* __restrict__ sys, const long long* __restrict__ tmp,
                                         const int* __restrict__ newSlot, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevSys& S = *sys;
    long long* f = S.ommForce + newSlot[i];
    const size_t st = (size_t)S.ommStride;
    f[0] = tmp[3 * (size_t)i];
    f[st] = tmp[3 * (size_t)i + 1];
    f[2 * st] = tmp[3 * (size_t)i + 2];
}


// body force and space-frame torque from the current atom forces
template <class Real>
__device__ inline void body_force_torque(const DevSys& S, Q4<double> q, int N, int atom0, V3<double>& F,
                                         V3<double>& tau) {
    const int dbase = atom0 - S.numFree;
    tau = {0.0, 0.0, 0.0};
    F = {0.0, 0.0, 0.0};
    for (int j = 0; j < N; j++) {
        V3<double> dl = Atmul(q, cvt<double>(ld3<Real>(S.d, S.dStride, dbase + j)));
        V3<double> f = cvt<double>(ld3<Real>(S.force, S.atomStride, atom0 + j));
        F = F + f;
        tau = tau + cross(dl, f);
    }
}

template <class Real>
__global__ void k_body_states(const DevSys* __restrict__ sys, const int* __restrict__ bodyN,
                              const int* __restrict__ bodyAtom0, int numBodies, double* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numBodies) return;
    const DevSys& S = *sys;
    int N = bodyN[k], atom0 = bodyAtom0[k];
    Q4<double> q = cvt<double>(ld4<Real>(S.q, k));
    Q4<Real> iI = ld4<Real>(S.iI, k);
    V3<double> invI = {(double)iI.a, (double)iI.b, (double)iI.c};
    double mass = 0.0;
    for (int j = 0; j < N; j++) mass += S.mass[atom0 + j];
    V3<double> F, tau;
    body_force_torque<Real>(S, q, N, atom0, F, tau);
    // take out the folded half-kick of the next step
    double halfDt = 0.5 * S.dt;
    V3<double> pcm = cvt<double>(ldp3<Real>(S.pcm, k)) - F * halfDt;
    V3<double> L = cvt<double>(ldp3<Real>(S.L, k)) - body_torque(q, tau, invI.z) * halfDt;
    Q4<double> pi = Bmul(q, L) * 2.0;
    Q4<double> T = Cmul(q, tau);
    V3<double> w = hadamard(invI, L);
    V3<double> vcm = pcm * (double)iI.d;
    double* o = out + 36 * (size_t)k;
    int p = 0;
    o[p++] = N;
    o[p++] = invI.z == 0.0 ? 5.0 : 6.0;
    o[p++] = mass;
    o[p++] = invI.x != 0.0 ? 1.0 / invI.x : 0.0;
    o[p++] = invI.y != 0.0 ? 1.0 / invI.y : 0.0;
    o[p++] = invI.z != 0.0 ? 1.0 / invI.z : 0.0;
    o[p++] = invI.x; o[p++] = invI.y; o[p++] = invI.z;
    const int st = S.bodyStride;
    o[p++] = S.rcm[k]; o[p++] = S.rcm[k + st]; o[p++] = S.rcm[k + 2 * st];
    o[p++] = pcm.x; o[p++] = pcm.y; o[p++] = pcm.z;
    o[p++] = q.a; o[p++] = q.b; o[p++] = q.c; o[p++] = q.d;
    o[p++] = pi.a; o[p++] = pi.b; o[p++] = pi.c; o[p++] = pi.d;
    o[p++] = F.x; o[p++] = F.y; o[p++] = F.z;
    o[p++] = T.a; o[p++] = T.b; o[p++] = T.c; o[p++] = T.d;
    o[p++] = dot(pcm, vcm);
    o[p++] = dot(L, w);
    o[p++] = 0.0;
    for (; p < 36; p++) o[p] = 0.0;
}

template <class Real>
__global__ void k_rekick(const DevSys* __restrict__ sys, const int* __restrict__ bodyN,
                         const int* __restrict__ bodyAtom0, int numBodies, double ddt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numBodies) return;
    const DevSys& S = *sys;
    Q4<double> q = cvt<double>(ld4<Real>(S.q, k));
    V3<double> F, tau;
    body_force_torque<Real>(S, q, bodyN[k], bodyAtom0[k], F, tau);
    V3<double> pcm = cvt<double>(ldp3<Real>(S.pcm, k)) + F * (0.5 * ddt);
    V3<double> L = cvt<double>(ldp3<Real>(S.L, k)) + body_torque(q, tau, (double)ld4<Real>(S.iI, k).c) * (0.5 * ddt);
    stp3<Real>(S.pcm, k, cvt<Real>(pcm));
    stp3<Real>(S.L, k, cvt<Real>(L));
}

// the next step's first half-kick, folded into the momenta with the forces
// `prev` (the layout of the force array), redone with the forces now in the
// force array: p += (F - F_prev) dt/2, L += A(q)(tau - tau_prev) dt/2; bit
// for bit unchanged where the forces are
template <class Real>
__global__ void k_refold(const DevSys* __restrict__ sys, const int* __restrict__ bodyN,
                         const int* __restrict__ bodyAtom0, int numBodies, const void* __restrict__ prev) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numBodies) return;
    const DevSys& S = *sys;
    const int N = bodyN[k], atom0 = bodyAtom0[k], dbase = atom0 - S.numFree;
    Q4<double> q = cvt<double>(ld4<Real>(S.q, k));
    V3<double> dF = {0.0, 0.0, 0.0}, dtau = {0.0, 0.0, 0.0};
    for (int j = 0; j < N; j++) {
        V3<double> dl = Atmul(q, cvt<double>(ld3<Real>(S.d, S.dStride, dbase + j)));
        V3<double> f = cvt<double>(ld3<Real>(S.force, S.atomStride, atom0 + j)) -
                       cvt<double>(ld3<Real>(prev, S.atomStride, atom0 + j));
        dF = dF + f;
        dtau = dtau + cross(dl, f);
    }
    const double halfDt = 0.5 * S.dt;
    V3<double> pcm = cvt<double>(ldp3<Real>(S.pcm, k)) + dF * halfDt;
    V3<double> L = cvt<double>(ldp3<Real>(S.L, k)) + body_torque(q, dtau, (double)ld4<Real>(S.iI, k).c) * halfDt;
    stp3<Real>(S.pcm, k, cvt<Real>(pcm));
    stp3<Real>(S.L, k, cvt<Real>(L));
}

// body-atom positions r_cm + A(q)^T d_j in FP64 into the staging buffer,
// internal atom a at row orig[a] (row a when orig is null); the rotation is
// the one part1 applies
template <class Real>
__global__ void k_body_atoms_out(const DevSys* __restrict__ sys, const int* __restrict__ bodyN,
                                 const int* __restrict__ bodyAtom0, int numBodies, double* __restrict__ out,
                                 const int* __restrict__ orig) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numBodies) return;
    const DevSys& S = *sys;
    const int st = S.bodyStride;
    V3<double> rc = {S.rcm[k], S.rcm[k + st], S.rcm[k + 2 * st]};
    Rot<Real> R = rotT(ld4<Real>(S.q, k));
    int N = bodyN[k], atom0 = bodyAtom0[k];
    for (int j = 0; j < N; j++) {
        V3<double> r = rc + cvt<double>(apply(R, ld3<Real>(S.d, S.dStride, atom0 + j - S.numFree)));
        size_t a = orig ? (size_t)orig[atom0 + j] : (size_t)(atom0 + j);
        out[3 * a] = r.x;
        out[3 * a + 1] = r.y;
        out[3 * a + 2] = r.z;
    }
}

// Bodies of one shape.  The bodies of a size class (count bodies of N atoms
// from firstBody, their d rows from row0) are compared with candidate shapes:
// the d rows of a few bodies of the class, given by their first row.  A body
// whose coordinates all lie within tol of a candidate's, and whose 1/I and
// 1/M lie within 2^-21 of the candidate's relative to them, takes that
// candidate's values into its own rows and inertia and the candidate's body
// index as DevSys::dOff, so the step kernels of a sharing tile read the
// shared rows and inertia (few, cached) and every other reader sees the same
// values in the body's own.  The bodies that matched are counted in
// *matched.  A candidate matches itself and copies nothing; a body that is
// no candidate writes only its own rows.
template <class Real>
__global__ void k_match_frames(const DevSys* __restrict__ sys, int firstBody, int count, int N, int row0,
                               const int* __restrict__ candRows, int numCand, double tol, int* matched) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    const DevSys& S = *sys;
    const int own = row0 + b * N, k = firstBody + b;
    const Q4<Real> ownI = ld4<Real>(S.iI, k);
    auto near = [](Real x, Real y) {
        const double a = (double)x, c = (double)y;
        return fabs(a - c) <= 0x1p-21 * fmax(fabs(a), fabs(c));
    };
    for (int c = 0; c < numCand; c++) {
        const int cr = candRows[c], kc = firstBody + (cr - row0) / N;
        const Q4<Real> candI = ld4<Real>(S.iI, kc);
        bool same = near(ownI.a, candI.a) && near(ownI.b, candI.b) && near(ownI.c, candI.c) && near(ownI.d, candI.d);
        for (int j = 0; same && j < N; j++) {
            V3<Real> x = ld3<Real>(S.d, S.dStride, own + j), y = ld3<Real>(S.d, S.dStride, cr + j);
            same = fabs((double)x.x - (double)y.x) <= tol && fabs((double)x.y - (double)y.y) <= tol &&
                   fabs((double)x.z - (double)y.z) <= tol;
        }
        if (same) {
            if (cr != own) {
                for (int j = 0; j < N; j++) st3<Real>(S.d, S.dStride, own + j, ld3<Real>(S.d, S.dStride, cr + j));
                st4<Real>(S.iI, k, candI);
            }
            S.dOff[k] = kc;
            atomicAdd(matched, 1);
            return;
        }
    }
}

// ---------------------------------------------------------------------------
hipError_t launch_match_frames(int precision, const DevSys* sys, int firstBody, int count, int N, int row0,
                               const int* candRows, int numCand, double tol, int* matched, hipStream_t stream) {
    if (count <= 0 || numCand <= 0) return hipSuccess;
    const int grid = (count + 255) / 256;
    if (precision == 0)
        k_match_frames<float><<<grid, 256, 0, stream>>>(sys, firstBody, count, N, row0, candRows, numCand, tol, matched);
    else
        k_match_frames<double><<<grid, 256, 0, stream>>>(sys, firstBody, count, N, row0, candRows, numCand, tol, matched);
    return hipGetLastError();
}

hipError_t launch_setup(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int geometry,
                        int block, hipStream_t stream) {
    if (numTiles <= 0) return hipSuccess;
    if (precision == 0) k_setup<float><<<numTiles, block, 0, stream>>>(systems, tiles, numTiles, geometry);
    else k_setup<double><<<numTiles, block, 0, stream>>>(systems, tiles, numTiles, geometry);
    return hipGetLastError();
}

hipError_t launch_stage_in(int precision, const DevSys* sys, const double* in, int numAtoms, int what,
                           hipStream_t stream, int adopted, const int* orig, double* keep) {
    if (numAtoms <= 0) return hipSuccess;
    if (adopted && what == 2) return hipErrorInvalidValue;
    int grid = (numAtoms + 255) / 256;
    if (precision == 0) {
        if (adopted) k_stage_in<float, AdoptedAtoms><<<grid, 256, 0, stream>>>(sys, in, numAtoms, what, orig, keep);
        else k_stage_in<float, SoaAtoms><<<grid, 256, 0, stream>>>(sys, in, numAtoms, what, orig, keep);
    } else {
        if (adopted) k_stage_in<double, AdoptedAtoms><<<grid, 256, 0, stream>>>(sys, in, numAtoms, what, orig, keep);
        else k_stage_in<double, SoaAtoms><<<grid, 256, 0, stream>>>(sys, in, numAtoms, what, orig, keep);
    }
    return hipGetLastError();
}

hipError_t launch_stage_out(int precision, const DevSys* sys, double* out, int numAtoms, int what,
                            hipStream_t stream, int adopted, const int* orig) {
    if (numAtoms <= 0) return hipSuccess;
    int grid = (numAtoms + 255) / 256;
    if (precision == 0) {
        if (adopted) k_stage_out<float, AdoptedAtoms><<<grid, 256, 0, stream>>>(sys, out, numAtoms, what, orig);
        else k_stage_out<float, SoaAtoms><<<grid, 256, 0, stream>>>(sys, out, numAtoms, what, orig);
    } else {
        if (adopted) k_stage_out<double, AdoptedAtoms><<<grid, 256, 0, stream>>>(sys, out, numAtoms, what, orig);
        else k_stage_out<double, SoaAtoms><<<grid, 256, 0, stream>>>(sys, out, numAtoms, what, orig);
    }
    return hipGetLastError();
}

hipError_t launch_tether_in(int precision, const DevSys* sys, const double* in, int numAtoms, int what,
                            const int* orig, hipStream_t stream) {
    if (numAtoms <= 0) return hipSuccess;
    int grid = (numAtoms + 255) / 256;
    if (precision == 0) k_tether_in<float><<<grid, 256, 0, stream>>>(sys, in, numAtoms, what, orig);
    else k_tether_in<double><<<grid, 256, 0, stream>>>(sys, in, numAtoms, what, orig);
    return hipGetLastError();
}

hipError_t launch_adopted_forces_in(int precision, const DevSys* sys, int numAtoms, hipStream_t stream) {
    if (numAtoms <= 0) return hipSuccess;
    int grid = (numAtoms + 255) / 256;
    if (precision == 0) k_adopted_forces_in<float><<<grid, 256, 0, stream>>>(sys, numAtoms);
    else k_adopted_forces_in<double><<<grid, 256, 0, stream>>>(sys, numAtoms);
    return hipGetLastError();
}

hipError_t launch_adopted_move_forces(const DevSys* sys, long long* tmp, const int* newSlot, int numAtoms,
                                      hipStream_t stream) {
    if (numAtoms <= 0) return hipSuccess;
    int grid = (numAtoms + 255) / 256;
    k_adopted_gather_forces<<<grid, 256, 0, stream>>>(sys, tmp, numAtoms);
    k_adopted_scatter_forces<<<grid, 256, 0, stream>>>(sys, tmp, newSlot, numAtoms);
    return hipGetLastError();
}

hipError_t launch_body_states(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0,
                              int numBodies, double* out, hipStream_t stream) {
    if (numBodies <= 0) return hipSuccess;
    int grid = (numBodies + 255) / 256;
    if (precision == 0) k_body_states<float><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, out);
    else k_body_states<double><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, out);
    return hipGetLastError();
}

hipError_t launch_body_atoms_out(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0,
                                int numBodies, double* out, hipStream_t stream, const int* orig) {
    if (numBodies <= 0) return hipSuccess;
    int grid = (numBodies + 255) / 256;
    if (precision == 0)
        k_body_atoms_out<float><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, out, orig);
    else k_body_atoms_out<double><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, out, orig);
    return hipGetLastError();
}

hipError_t launch_refold(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0, int numBodies,
                         const void* prev, hipStream_t stream) {
    if (numBodies <= 0) return hipSuccess;
    int grid = (numBodies + 255) / 256;
    if (precision == 0) k_refold<float><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, prev);
    else k_refold<double><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, prev);
    return hipGetLastError();
}

hipError_t launch_rekick(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0, int numBodies,
                         double ddt, hipStream_t stream) {
    if (numBodies <= 0) return hipSuccess;
    int grid = (numBodies + 255) / 256;
    if (precision == 0) k_rekick<float><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, ddt);
    else k_rekick<double><<<grid, 256, 0, stream>>>(sys, bodyN, bodyAtom0, numBodies, ddt);
    return hipGetLastError();
}

}  // namespace rbhip
