// Shared host/device layout of the rigid-body step library (gfx950).
//
// Device data layout (structure of arrays, one context = one system):
//   bodies, internal body order (grouped by atoms-per-body class):
//     q              Real4[numBodies]   orientation quaternion
//     L              Real[3*numBodies]  body-frame angular momentum, xyz per body
//                                       (the reference's pi = 2 B(q) L)
//     pcm            Real[3*numBodies]  centre-of-mass momentum, xyz per body
//                                       (L and pcm carry the next step's first
//                                       half-kick, folded into part2)
//     iI             Real4[numBodies]   (1/I0, 1/I1, 1/I2, 1/M); 1/I2 = 0 marks a
//                                       linear (dof 5) body
//     rcm            double[3][bodyStride]  centre of mass, FP64
//     d              Real[3][dStride]   body-fixed atom coordinates
//   atoms, internal order (free atoms, then each body's atoms contiguously),
//   one array per component (x[], y[], z[], row length atomStride):
//     pos            Real[3][atomStride]   positions (hi part in mixed precision)
//     posLo          float[3][atomStride]  low part, mixed precision, free atoms
//                    (a body atom's exact position is r_cm + A(q)^T d; the
//                    force field reads its fp32 value, as OpenMM's mixed
//                    precision force kernels read fp32 posq)
//     vel, force     Real[3][atomStride]
//     freeInvM       Real[numFree]
// With a lane per body, the N atoms of a body are N consecutive elements of
// each component array, so a wave's loads and stores are contiguous runs.
// Work is described by tiles: each workgroup iteration handles one tile of
// free atoms, of bodies of one size class, or of atoms for the force kernel.
#pragma once

#include <cstdint>

namespace rbhip {

// virtual-site kinds (the values of include/rbhip.h's RBHIP_SITE_*)
enum SiteKind : int { SITE_TWO_AVERAGE = 1, SITE_THREE_AVERAGE = 2, SITE_OUT_OF_PLANE = 3, SITE_LOCAL_COORDINATES = 4 };

enum TileKind : int { TILE_FREE = 0, TILE_BODIES = 1, TILE_ATOMS = 2, TILE_CLUSTERS = 3 };

// free-atom tiles cover this many atoms per workgroup thread (the integrator
// kernels loop over them), so the per-tile work of the kinetic-energy
// partials is spread over more atoms
constexpr int FREE_PER_THREAD = 4;
// atoms per workgroup thread in the force kernel's tiles
constexpr int FORCE_PER_THREAD = 4;

struct Tile {
    int sys;        // index into the system descriptor table
    int kind;       // TileKind
    int first;      // first body (TILE_BODIES) or first atom (TILE_FREE / TILE_ATOMS)
    int count;      // bodies or atoms in the tile
    int N;          // atoms per body (TILE_BODIES)
    int atomFirst;  // internal index of the tile's first body atom (TILE_BODIES)
    int lanes;      // lanes per body, a power of two (TILE_BODIES)
    int slot;       // reduction slot of this tile in the system's partials
    // 1: the tile's bodies read their body-fixed coordinates and inertia at
    // the body DevSys::dOff names (bodies of one shape share them), 0: their own
    int sharedD = 0;
};

// One system as the kernels see it.  Pointers to Real data are void* and
// cast by the kernels (Real = float for mixed, double for double precision).
struct DevSys {
    int numFree, numBodies, numBodyAtoms, numAtoms;
    int rotationMode;   // 0 exact, n > 0 NO-SQUISH(n)
    int bodyStride;     // row length of rcm
    int useStageR, useStageV;
    int atomStride;     // row length of the per-atom component arrays
    int dStride;        // row length of d
    double dt;
    double tetherScale; // 1 if the tether is on, 0 otherwise
    int tetherUniform;  // 1: every atom has spring constant tetherK0 (tetherK unused)
    int pad0;
    double tetherK0;
    // bodies
    void* q;
    void* L;
    void* pcm;
    void* iI;
    double* rcm;
    void* d;
    // per body, the body whose d rows and inertia (1/I, 1/M) it reads when
    // its tile shares them: itself, or a body of the same size class whose
    // coordinates and inertia it equals to within a few fp32 ulps (its own
    // rows and inertia then hold the same values)
    int* dOff;
    // atoms
    void* pos;
    float* posLo;
    void* vel;
    void* force;
    void* freeInvM;
    double* mass;
    double* stageR;     // FP64 xyz-interleaved staging, internal order
    double* stageV;
    // force field
    void* tetherK;      // Real[numAtoms]
    void* r0;           // Real[3][atomStride], the state precision
    int* bondStart;     // [numAtoms + 1] CSR of bonds per atom (nullptr: none)
    int* bondOther;     // partner atom of each bond entry
    double* bondLen;
    double* bondK;
    // reductions: 4 doubles per slot
    double* kePart;     // twoKt, twoKr, dof, (unused)
    double* pePart;     // energy, (unused) x3
    // refined energies (all nullptr unless enabled)
    void* refBody;      // Real4[2 * numBodies]: (a_r, 0), a_q per body, written at the
                        // start of a call's last step, read at its end
    void* refFree;      // Real[3 * numFree] packed: a_v per free atom, likewise
    double* refKePart;  // per part tile: 6 dt KE_t refined, 6 dt KE_r refined
    double* refPePart;  // per part tile: sum |F|^2/m + tau.(I^-1 tau)
    // distance constraints between free atoms (numClusters = 0 without): the
    // atoms of each connected cluster of constraints are contiguous, after
    // the unconstrained free atoms
    int numClusters;
    int pad1;
    int* clusterStart;  // [numClusters + 1] first internal atom of each cluster
    int* consStart;     // [numClusters + 1] first constraint of each cluster
    int* consA;         // internal atom indices of each constraint, in the order given
    int* consB;
    double* consDist;
    double constraintTol;
    unsigned* constraintFail;   // count of cluster iterations that did not converge
    // virtual sites on massless particles (numSites = 0 without).  Massless
    // particles follow the body atoms in internal order and are never
    // integrated.  Each parent atom of a site gathers the force of its sites
    // in a fixed order: parentAtom[g], and the (site, role) entries
    // [parentStart[g], parentStart[g + 1]) of parentSite / parentRole.
    int numSites;
    int numSiteParents;
    int* siteAtom;      // internal index of each site
    int* siteType;      // RBHIP_SITE_*
    int* siteParent;    // 3 internal parent indices per site
    double* siteW;      // 12 weights per site (3 used but by local-coordinates sites)
    int* parentAtom;
    int* parentStart;
    int* parentSite;
    int* parentRole;    // 0, 1, 2: which parent of the site
    // Atom state adopted from a caller in OpenMM's GPU layout (ommSlot ==
    // nullptr: the library's own arrays above hold it).  Per atom slot s:
    //   ommPosq      Real4[s] = (x, y, z, charge)         charge left alone
    //   ommPosqCorr  float4[s] = low parts of x, y, z     mixed precision only
    //   ommVelm      double4[s] = (vx, vy, vz, 1/m)        1/m left alone
    //   ommForce     long long[3][ommStride], fixed point with 2^32 per unit
    // The step kernels reach internal atom i at slot ommSlot[i].
    int* ommSlot;
    void* ommPosq;
    float* ommPosqCorr;
    double* ommVelm;
    long long* ommForce;
    int ommStride;
    // Constraints applied by the caller to adopted arrays (OpenMM's
    // applyConstraints between the free atoms' drift and the rest of part 1):
    // posDelta (double4 per slot) holds the constrained displacements during
    // a constrained part 1, which keeps each free atom's correction (the
    // constrained minus the unconstrained drift, Real[3][numFree]) in
    // freeCorr; corrPending = 1 makes the next part 2 add correction / dt to
    // the free atoms' velocities, as the reference's (R - saved) / dt does
    // (RigidBodySystem.cpp:196-197).
    int corrPending;
    double* posDelta;
    void* freeCorr;
};

// OpenMM's fixed-point force unit: 2^32 per kJ/mol/nm
constexpr double kFixedPointScale = 4294967296.0;

}  // namespace rbhip
