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

}  // namespace rbhip
