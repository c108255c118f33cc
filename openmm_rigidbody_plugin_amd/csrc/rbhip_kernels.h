// Kernel launchers shared by the host side of the library.
#pragma once

#include <hip/hip_runtime.h>

#include "rbhip_internal.h"

namespace rbhip {

// events for the next step kernel (k_part1, k_part2, k_fused or k_forces)
// launched on this thread: attached to its dispatch, they record its own
// start and end; cleared by that launch
void set_step_kernel_events(hipEvent_t start, hipEvent_t stop);
// clears them; true if no step kernel took them
bool clear_step_kernel_events();

// hot path (rbhip_kernels.hip); `adopted` = the atoms live in adopted
// arrays in OpenMM's layout (DevSys::ommSlot), else in the library's own
hipError_t launch_part1(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                        int block, int maxBlocks, hipStream_t stream, int adopted);
hipError_t launch_part2(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                        int maxBlocks, hipStream_t stream, int adopted);
// adopted arrays: the free atoms' drift displacements into the caller's
// posDelta (double4 per slot) for its constraint solver
hipError_t launch_free_delta(int precision, const DevSys* sys, int numFree, double* posDelta, hipStream_t stream);
// adopted arrays: the free-atom tiles' KE partials from the velocities as
// they are now (after the caller's velocity constraints)
hipError_t launch_free_ke(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                          hipStream_t stream);
// part2 of step n and part1 of step n+1 in one pass (no atom velocities, no KE)
hipError_t launch_fused(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                        int block, int maxBlocks, hipStream_t stream, int adopted);
// refined energies of the systems whose refBody is set: the start terms of
// the next step (`folded` = the stored momenta carry its first half-kick,
// i.e. part1 follows; 0 = a fused boundary follows), the end terms after
// its part2, and the potential refinement of the current forces
hipError_t launch_refine_begin(int precision, int rotationMode, const DevSys* systems, const Tile* tiles,
                               int numTiles, int folded, int block, int maxBlocks, hipStream_t stream, int adopted);
hipError_t launch_refine_end(int precision, int rotationMode, const DevSys* systems, const Tile* tiles, int numTiles,
                             int block, int maxBlocks, hipStream_t stream, int adopted);
hipError_t launch_refine_pe(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                            int maxBlocks, hipStream_t stream, int adopted);
// (energy: also the tiles' potential-energy partials)
hipError_t launch_forces(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int block,
                         int maxBlocks, hipStream_t stream, int adopted, int energy);
// virtual sites of one system: positions from the parents; after the forces,
// the sites' forces moved to their parents and cleared
hipError_t launch_site_positions(int precision, const DevSys* sys, int numSites, hipStream_t stream);
hipError_t launch_site_forces(int precision, const DevSys* sys, int numSites, int numParents, hipStream_t stream);
// clusters of constrained free atoms (TILE_CLUSTERS tiles, one lane per
// cluster, block <= 256): part 1 (kick, drift, SHAKE), part 2 (kick, RATTLE,
// KE partials), both at a fused boundary, or the KE partials after the set-up
enum ClusterPhase : int { CLUSTER_PART1 = 1, CLUSTER_PART2 = 2, CLUSTER_FUSED = 3, CLUSTER_KE = 4 };
constexpr int CLUSTER_MAX_ATOMS = 16;
constexpr int CLUSTER_MAX_ITERATIONS = 1000;
// (counts: tiles of clusters of n = 2, 3 and 4 atoms with at most
// n (n - 1) / 2 constraints, then of the other clusters, in that order)
hipError_t launch_clusters(int precision, const DevSys* systems, const Tile* tiles, const int* counts, int phase,
                           int block, hipStream_t stream);
// which: 0 kePart, 1 pePart, 2 refKePart, 3 refPePart
hipError_t launch_reduce(const DevSys* systems, const int* nslots, int numSystems, int which, double* out,
                         hipStream_t stream);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t stream);

// setup and state conversion (rbhip_setup.hip, FP64 without contraction)
// builds the body state (geometry if `geometry`, always dynamics) from the
// staged or current atom state, records KE and dof partials, and applies the
// first half-kick of the next step
hipError_t launch_setup(int precision, const DevSys* systems, const Tile* tiles, int numTiles, int geometry,
                        int block, hipStream_t stream);
// bodies of one size class whose body-fixed coordinates equal a candidate's
// to within tol, and whose inertia equals it to within 2^-21, take its values
// and share its d rows and inertia (DevSys::dOff); the count of such bodies
// is added to *matched (rbhip_setup.hip)
hipError_t launch_match_frames(int precision, const DevSys* sys, int firstBody, int count, int N, int row0,
                               const int* candRows, int numCand, double tol, int* matched, hipStream_t stream);
// FP64 xyz rows -> device state: what = 0 positions, 1 velocities, 2 forces
// (the library's force array; not with `adopted`).  Internal atom i reads row
// orig[i] (row i when orig is null); with `keep`, its FP64 row is also left
// at row i of keep (internal order).
hipError_t launch_stage_in(int precision, const DevSys* sys, const double* in, int numAtoms, int what,
                           hipStream_t stream, int adopted, const int* orig = nullptr, double* keep = nullptr);
// device state -> FP64 xyz rows: what = 0 positions, 1 velocities, 2 forces;
// internal atom i goes to row orig[i] (row i when orig is null)
hipError_t launch_stage_out(int precision, const DevSys* sys, double* out, int numAtoms, int what,
                            hipStream_t stream, int adopted, const int* orig = nullptr);
// FP64 tether input in original order -> the state-precision arrays: what = 0
// anchors (xyz rows), 1 spring constants; internal atom i reads row orig[i]
hipError_t launch_tether_in(int precision, const DevSys* sys, const double* in, int numAtoms, int what,
                            const int* orig, hipStream_t stream);
// adopted fixed-point forces -> the library's force array (state precision)
hipError_t launch_adopted_forces_in(int precision, const DevSys* sys, int numAtoms, hipStream_t stream);
// moves the adopted forces of the actual atoms from the slots of DevSys::ommSlot
// to newSlot (tmp: 3 * numAtoms long longs)
hipError_t launch_adopted_move_forces(const DevSys* sys, long long* tmp, const int* newSlot, int numAtoms,
                                      hipStream_t stream);
// body states (internal body order) as 36 FP64 values each, half-kick removed
hipError_t launch_body_states(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0,
                              int numBodies, double* out, hipStream_t stream);
// FP64 body-atom positions rebuilt from the body state into `out`, internal
// atom a at row orig[a] (row a when orig is null)
hipError_t launch_body_atoms_out(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0,
                                 int numBodies, double* out, hipStream_t stream, const int* orig = nullptr);
// redoes the folded half-kick for the forces now in the force array, made
// before with the forces `prev` (a copy of the force array)
hipError_t launch_refold(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0, int numBodies,
                         const void* prev, hipStream_t stream);
// re-applies the folded half-kick for a new step size: ddt = dtNew - dtOld
hipError_t launch_rekick(int precision, const DevSys* sys, const int* bodyN, const int* bodyAtom0, int numBodies,
                         double ddt, hipStream_t stream);

}  // namespace rbhip
