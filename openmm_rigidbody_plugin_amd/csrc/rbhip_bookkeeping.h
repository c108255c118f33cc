// Device-side bookkeeping of the rigid-body system (rbhip_bookkeeping.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace rbhip {

// scratch bytes launch_bookkeeping needs for numAtoms atoms
hipError_t bookkeeping_scratch_bytes(int numAtoms, hipStream_t stream, size_t* bytes);

// cleanBodyIndices and the atom order of RigidBodySystem::initialize for n
// atoms (device arrays): bodyIndex[n], atomIndex[free + body atoms],
// bodyAtomStart[numBodies + 1] (each body's first atom counted from the
// first body atom; the last entry is the number of body atoms), using
// `scratch` (scratchBytes, 256-byte aligned).  Queued on `stream`, which it
// then waits for; smallOut (host, 5 ints) receives the status flags (1: a
// massless particle in a body, 2: a negative or invalid mass) and numFree,
// numBodies, numMassless, numBodyAtoms.
hipError_t launch_bookkeeping(int n, const int* labels, const double* masses, int* bodyIndex, int* atomIndex,
                              int* bodyAtomStart, void* scratch, size_t scratchBytes, hipStream_t stream,
                              int* smallOut);

// The internal atom order of a context, atomOrig[internal atom] = original
// atom: the numFree free atoms first (freeOrder, or the reference's free
// atoms atomIndex[0, numFree) when freeOrder is null), then the bodies in
// internal order (internal body ib is reference body order[ib]; its bodyN[ib]
// atoms, from bodyAtom0[ib] on, are the reference body's atoms in the
// reference's order, atomIndex[numFree + bodyAtomStart[order[ib]] + j]),
// then the numMassless massless particles at massStart (massless[k]).
// Device arrays; queued on `stream`.
hipError_t launch_internal_order(int numFree, int numBodies, int numMassless, int massStart, const int* atomIndex,
                                 const int* bodyAtomStart, const int* freeOrder, const int* order, const int* bodyN,
                                 const int* bodyAtom0, const int* massless, int* atomOrig, hipStream_t stream);

// masses in internal order, mass[i] = raw[atomOrig[i]], and the free atoms'
// inverse masses at the state precision (precision 0: fp32, else FP64)
hipError_t launch_masses_in(int precision, int n, int numFree, const int* atomOrig, const double* raw, double* mass,
                            void* freeInvM, hipStream_t stream);

}  // namespace rbhip
