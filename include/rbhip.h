/*
 * rbhip.h — C ABI of the MI355X (gfx950) RigidBodyIntegrator step library.
 *
 * This is the boundary between the host C++ plugin layer (a drop-in
 * HipIntegrateRigidBodyStepKernel behind the reference's unchanged
 * RigidBodyIntegrator / IntegrateRigidBodyStepKernel API) and the
 * hand-written HIP kernels.  Plain pointers and sizes, int status codes, a
 * thread-local last-error string, no exceptions across the ABI.
 *
 * Each entry point cites the reference interface it replaces (paths relative
 * to the reference repository).
 *
 * Conventions
 *   - Host arrays of vectors are xyz-interleaved FP64 in ORIGINAL atom order.
 *   - Status: RBHIP_OK (0) or a negative RBHIP_E_* code; rbhip_last_error()
 *     describes the last failure on the calling thread.
 *   - One rbhip_context = one simulated system on one device: the platform's
 *     atom arrays (positions, velocities, forces), its force field and the
 *     rigid-body integrator state.  A context is driven by one host thread.
 *   - Device memory holds atoms in an internal order (free atoms, then body
 *     atoms grouped by body size class, each body's atoms contiguous); the
 *     get/set entry points translate to and from the original order.
 */
#ifndef RBHIP_H_
#define RBHIP_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBHIP_VERSION 1

enum {
    RBHIP_OK = 0,
    RBHIP_E_ARG = -1,       /* invalid argument */
    RBHIP_E_HIP = -2,       /* HIP runtime failure */
    RBHIP_E_STATE = -3,     /* call not valid in the current state */
    RBHIP_E_NOMEM = -4,     /* device allocation failed */
    RBHIP_E_UNSUPPORTED = -5
};

enum { RBHIP_PRECISION_MIXED = 0, RBHIP_PRECISION_DOUBLE = 1 };

/* indices of the per-kernel timing array filled by rbhip_time_kernels */
enum {
    RBHIP_KERNEL_PART1 = 0,
    RBHIP_KERNEL_FORCES = 1,
    RBHIP_KERNEL_PART2 = 2,
    RBHIP_KERNEL_FUSED = 3, /* part2 of step n + part1 of step n+1 in one pass */
    RBHIP_NUM_KERNELS = 4
};

typedef struct rbhip_context rbhip_context;
typedef struct rbhip_batch rbhip_batch;
typedef void* rbhip_stream; /* a hipStream_t; NULL = the library's own stream */

const char* rbhip_last_error(void);
int rbhip_version(void);
/* number of visible devices (0 without a GPU) */
int rbhip_device_count(int* count);
/* hipDeviceSynchronize on the current device */
int rbhip_device_synchronize(void);

/* ---------------------------------------------------------------------- */
/* Context lifetime                                                        */
/* ---------------------------------------------------------------------- */
/* Creates the device context and runs the integrator's initialize:
 *   RigidBodyIntegrator::initialize   (openmmapi/src/RigidBodyIntegrator.cpp:40-51)
 *   RigidBodySystem::initialize       (openmmapi/src/RigidBodySystem.cpp:55-114)
 *   IntegrateRigidBodyStepKernel::initialize (openmmapi/include/RigidBodyKernels.h:60)
 * bodyIndices follow the reference convention (<= 0 free atom, equal positive
 * labels one rigid body).  rotationMode: 0 exact, n > 0 NO-SQUISH with n
 * substeps (RigidBodyIntegrator.h:58-63).  Errors mirror the reference's
 * exceptions (negative mode, constraints on body atoms) plus this platform's
 * documented limits (single-atom bodies, massless particles in a body). */
int rbhip_create(int numAtoms, const double* masses, const int* bodyIndices, double stepSize,
                 int rotationMode, int precision, int device, rbhip_context** out);
/* rbhip_create for a System with distance constraints (System::addConstraint):
   constraintAtoms holds two atom indices per constraint, constraintDistances
   one length each.  As the reference kernel does with OpenMM's constraints
   (ReferenceRigidBodyKernels.cpp:92-104), every step projects the free atoms'
   positions after part 1 and their velocities after part 2; constraints on
   body atoms are refused with the reference's message
   (RigidBodySystem.cpp:107-112) and each constraint removes a degree of
   freedom (RigidBodySystem.cpp:130).  The solver is SHAKE and RATTLE's
   velocity stage over each connected cluster of constraints, in FP64; a
   cluster may join at most 16 atoms.  Contexts with constraints cannot join
   a batch and do not compute refined energies. */
int rbhip_create_constrained(int numAtoms, const double* masses, const int* bodyIndices, int numConstraints,
                             const int* constraintAtoms, const double* constraintDistances, double stepSize,
                             int rotationMode, int precision, int device, rbhip_context** out);
/* OpenMM virtual sites on massless particles (rbhip_create accepts massless
   particles labelled free: like the reference, RigidBodySystem.cpp:66-76, it
   integrates neither them nor sites).  Each site is recomputed from its
   parents after part 1 and the constraints, as ReferenceRigidBodyKernels.cpp
   :100 does, and after every force evaluation its force moves to its parents
   (the transpose of d r_site / d r_parent) and is cleared.  Per site: a type,
   the site particle, 3 parents (the third unused for two-particle averages)
   and 12 weights, of which a site uses {w1, w2} (two-particle average),
   {w1, w2, w3} (three-particle average), {w12, w13, wCross} (out of plane:
   r = r1 + w12 r12 + w13 r13 + wCross r12 x r13) or all 12 (local
   coordinates: origin, x and y weights of the 3 parents, then the local
   position; r = o + p_x e_x + p_y e_y + p_z e_z with e_x along the x-weighted
   sum, e_z along (x-sum) x (y-sum), e_y = e_z x e_x).  Parents must be
   massive.  Set once, before stepping; contexts with sites cannot join a
   batch or compute refined energies. */
enum { RBHIP_SITE_TWO_AVERAGE = 1, RBHIP_SITE_THREE_AVERAGE = 2, RBHIP_SITE_OUT_OF_PLANE = 3,
       RBHIP_SITE_LOCAL_COORDINATES = 4 };
int rbhip_set_virtual_sites(rbhip_context* ctx, int count, const int* types, const int* siteAtoms,
                            const int* parents, const double* weights);
/* Integrator::setConstraintTolerance (default 1e-5): positions converge when
   every constrained distance is within (1 +- tolerance) of its length,
   velocities when every bond's relative velocity is within tolerance of
   orthogonal to the bond */
int rbhip_set_constraint_tolerance(rbhip_context* ctx, double tolerance);
/* cluster projections that did not converge within 1000 sweeps, so far */
int rbhip_constraint_failures(rbhip_context* ctx, long long* count);
int rbhip_destroy(rbhip_context* ctx);
/* launch stream for all work of this context (hipStream_t; NULL restores the
   stream the library assigned, one of a per-device pool of non-blocking
   streams shared by contexts and batches).  Work queued on the previous
   stream completes before work on the new one; the call does not wait for
   it, unless the context holds captured graphs (then it waits for the old
   stream and drops them). */
int rbhip_set_stream(rbhip_context* ctx, rbhip_stream stream);
/* device bytes held by the context (for admission against an HBM budget) */
int rbhip_device_bytes(rbhip_context* ctx, size_t* bytes);

/* ---------------------------------------------------------------------- */
/* Integer bookkeeping (bit-exact contract with RigidBodySystem)           */
/* ---------------------------------------------------------------------- */
/* out5 = numFree, numBodies, numActualAtoms, numBodyAtoms, numDOF
   (RigidBodySystem.h:36-40; numDOF valid after the first set_positions) */
int rbhip_get_counts(rbhip_context* ctx, int* out5);
/* RigidBodySystem::atomIndex: free atoms ascending, then body 1's atoms,
   body 2's, ... (RigidBodySystem.cpp:84-105); out has numActualAtoms ints */
int rbhip_get_atom_index(rbhip_context* ctx, int* out);
/* cleanBodyIndices result (RigidBodySystem.cpp:28-49); out has numAtoms ints */
int rbhip_get_body_index(rbhip_context* ctx, int* out);
/* The same bookkeeping on the device, for labels and masses already in
   device memory: cleanBodyIndices (RigidBodySystem.cpp:28-49) and the atom
   order and body layout of RigidBodySystem::initialize (:55-105).  Inputs
   bodyIndices[numAtoms] and masses[numAtoms] and outputs are device arrays:
   bodyIndexOut[numAtoms] (0 free, 1..B by ascending label); atomIndexOut[F +
   body atoms] (free atoms ascending, then body 1's atoms ascending, body
   2's, ...); bodyAtomStartOut[B + 1] (room for numAtoms + 1), body b's first
   atom among the body atoms (its loc) and the body-atom count last.
   counts4 (host) = numFree, numBodies, numMassless, numBodyAtoms.  Massless
   particles are neither free nor body atoms (RigidBodySystem.cpp:66-76) and
   are refused with a body label, as rbhip_create refuses them.  Work is
   queued on `stream` (NULL: a library stream of `device`), which the call
   waits for; its scratch is allocated and freed in stream order. */
int rbhip_body_bookkeeping(int numAtoms, const int* bodyIndices, const double* masses, int* bodyIndexOut,
                           int* atomIndexOut, int* bodyAtomStartOut, int* counts4, int device, rbhip_stream stream);
/* No counterpart in the reference: rbhip_body_bookkeeping for callers that
   cannot allocate device memory (code under graph capture, serving loops,
   harnesses that forbid hipMalloc).  Same arguments and outputs; its scratch
   is a static 40 MiB device array per GPU, held until the process ends, so
   it allocates nothing.  Systems whose sort and scan need more (beyond about
   1.2 million atoms) are refused with RBHIP_E_ARG.  Calls on one GPU share
   that array: the library serialises them, and each waits for its stream
   (NULL: the device's null stream) before it returns; callers that queue
   other work into that scratch's lifetime from other threads race. */
int rbhip_body_bookkeeping_noalloc(int numAtoms, const int* bodyIndices, const double* masses, int* bodyIndexOut,
                                   int* atomIndexOut, int* bodyAtomStartOut, int* counts4, int device,
                                   rbhip_stream stream);

/* ---------------------------------------------------------------------- */
/* Force field slot (ContextImpl::calcForcesAndEnergy,                      */
/* ReferenceRigidBodyKernels.cpp:101 / CudaRigidBodyKernels.cpp:424)       */
/* ---------------------------------------------------------------------- */
/* harmonic tether F_i = -k_i (r_i - r0_i); k has numAtoms entries, r0 is xyz */
int rbhip_set_tether(rbhip_context* ctx, const double* k, const double* r0);
/* harmonic bonds E = k/2 (|r_b - r_a| - length)^2 (OpenMM HarmonicBondForce) */
int rbhip_add_bonds(rbhip_context* ctx, int count, const int* atomPairs, const double* length,
                    const double* k);
/* evaluates forces (and, if energy != NULL, the potential energy) */
int rbhip_compute_forces(rbhip_context* ctx, double* energy);
/* Forces computed outside the library (FP64, xyz per atom, original order),
   as the context's force evaluation leaves them for integratePart2
   (ReferenceRigidBodyKernels.cpp:101-102): a caller with its own force field
   steps with rbhip_part1, then its forces at the new positions through this
   call, then rbhip_part2.  Set after a state change or rbhip_part2 (when the
   bodies' momenta already carry the next step's first half-kick, made with
   the forces held before), the half-kick is redone with the new forces.
   Forces on virtual sites move to their parents.  rbhip_set_positions and
   rbhip_execute evaluate the library's own force field and replace them.
   Not for contexts with adopted arrays (their forces are in the array). */
int rbhip_set_forces(rbhip_context* ctx, const double* xyz);

/* ---------------------------------------------------------------------- */
/* Adopted atom arrays (OpenMM's GPU layout)                               */
/* ---------------------------------------------------------------------- */
/* The context keeps its atoms in a caller's device arrays in the layout of
   OpenMM's CUDA/HIP platforms, which the reference's CUDA kernels step in
   place (CudaRigidBodyKernels.cpp:377-444, rigidbodyintegrator.cu:270-396):
     posq            Real4[paddedNumAtoms]: x, y, z, charge (float4 in mixed
                     precision, double4 in double)
     posqCorrection  float4[paddedNumAtoms], low parts of x, y, z (mixed
                     precision; NULL in double)
     velm            double4[paddedNumAtoms]: vx, vy, vz, 1/m
     force           long long[3 * paddedNumAtoms], component-major, fixed
                     point with 2^32 per kJ/mol/nm
   atomOrder[s] (orderLength entries, a permutation of 0 .. orderLength-1)
   is the original atom held at slot s: OpenMM's getAtomIndex().  The step
   kernels read and write the arrays through the atom -> slot map (the
   reference's atomLocation, CudaRigidBodyKernels.cpp:258-265); charges and
   inverse masses (w) are never written.  Afterwards the state is read from
   the arrays at rbhip_state_changed; rbhip_part1 / rbhip_part2 step them,
   with the caller's force evaluation in between, and rbhip_execute steps
   them with the library's force field.  The get/set calls reach the arrays.
   Not with the library's own constraints or virtual sites (the caller's:
   rbhip_part1_constrained below; sites are the caller's to place after
   part 1) or batches; refined energies read the arrays too, but not with
   the caller's constraints.  The arrays must outlive the context or a
   later rbhip_destroy. */
int rbhip_adopt_arrays(rbhip_context* ctx, void* posq, void* posqCorrection, void* velm, long long* force,
                       int paddedNumAtoms, const int* atomOrder, int orderLength);
/* the arrays' atom order changed (ReorderListener::execute,
   CudaRigidBodyKernels.cpp:69-113): positions and velocities were permuted
   by their owner; the forces of the actual atoms move to the new slots
   here, as the reference's listener moves them */
int rbhip_set_atom_order(rbhip_context* ctx, const int* atomOrder, int orderLength);
/* RigidBodyIntegrator::stateChanged (RigidBodyIntegrator.cpp:63-74) for
   adopted arrays: after the caller wrote positions (and evaluated the
   forces there) or velocities into them, rebuilds the body state from the
   arrays (RigidBodySystem::update, RigidBodySystem.cpp:120-142) */
enum { RBHIP_STATE_POSITIONS = 1, RBHIP_STATE_VELOCITIES = 2 };
int rbhip_state_changed(rbhip_context* ctx, int what);
/* Constraints between free atoms applied by the caller to adopted arrays,
   the CUDA platform's order (CudaRigidBodyKernels.cpp:414-418, 429):
     rbhip_free_delta         the free atoms' drift displacements
                              (v + F (1/m) dt/2) dt into posDelta (OpenMM's
                              integration posDelta: double4 per slot, w = 0;
                              rigidbodyintegrator.cu:270-279);
     (the caller's applyConstraints adjusts posDelta)
     rbhip_part1_constrained  part 1 with the free atoms moved by the
                              constrained posDelta; the difference from the
                              unconstrained drift is kept, and the next
                              rbhip_part2 adds it / dt to their velocities,
                              as the reference's (R - saved) / dt
                              (RigidBodySystem.cpp:196-197,
                              ReferenceRigidBodyKernels.cpp:92-104);
     (forces, rbhip_part2, then the caller's applyVelocityConstraints)
   After such steps the kinetic energies re-read the free atoms' velocities
   from the arrays, so they include the caller's velocity constraints.  Not
   between rbhip_part1_constrained and rbhip_part2: another part 1,
   rbhip_free_delta, rbhip_execute, rbhip_time_kernels or rbhip_autotune;
   rbhip_state_changed abandons the pending correction.  posDelta is indexed
   by slot, so the atom order must not change between rbhip_free_delta and
   rbhip_part1_constrained (OpenMM reorders at the end of a step,
   CudaRigidBodyKernels.cpp:443). */
int rbhip_free_delta(rbhip_context* ctx, double* posDelta);
int rbhip_part1_constrained(rbhip_context* ctx, const double* posDelta);

/* ---------------------------------------------------------------------- */
/* Context state                                                           */
/* ---------------------------------------------------------------------- */
/* Context::setPositions -> RigidBodyIntegrator::stateChanged(Positions)
   (RigidBodyIntegrator.cpp:63-74): forces, then RigidBodySystem::update with
   geometry and velocities (RigidBodySystem.cpp:120-142), then
   uploadBodySystem (RigidBodyKernels.h:66).  The body state is built on the
   device in FP64. */
int rbhip_set_positions(rbhip_context* ctx, const double* xyz);
/* Context::setVelocities -> stateChanged(Velocities): dynamics rebuild */
int rbhip_set_velocities(rbhip_context* ctx, const double* xyz);
int rbhip_get_positions(rbhip_context* ctx, double* xyz);
int rbhip_get_velocities(rbhip_context* ctx, double* xyz);
int rbhip_get_forces(rbhip_context* ctx, double* xyz);
/* Integrator::setStepSize */
int rbhip_set_step_size(rbhip_context* ctx, double dt);
/* ContextImpl time and step count (ReferenceRigidBodyKernels.cpp:106-107) */
int rbhip_get_time(rbhip_context* ctx, double* time, long long* stepCount);

/* ---------------------------------------------------------------------- */
/* The step (IntegrateRigidBodyStepKernel::execute, RigidBodyKernels.h:73) */
/* ---------------------------------------------------------------------- */
/* integratePart1 (RigidBodySystem.cpp:170-187) on the device */
int rbhip_part1(rbhip_context* ctx);
/* integratePart2 (RigidBodySystem.cpp:193-204) on the device */
int rbhip_part2(rbhip_context* ctx);
/* RigidBodyIntegrator::step(steps) (RigidBodyIntegrator.cpp:96-101).
   With fusion (default) the interior step boundaries run as one pass each:
   part1, forces, (steps-1) x [fused part2+part1, forces], part2 -- the same
   trajectory as steps x [part1, forces, part2] (bit for bit in mixed
   precision, to a few ulp in FP64), without the interior steps' atom
   velocities and kinetic energies, which nothing reads.
   Repeated sequences are replayed from HIP graphs when enabled. */
int rbhip_execute(rbhip_context* ctx, int steps);
/* 1 = capture the repeated launches in HIP graphs and replay them (default) */
int rbhip_use_graph(rbhip_context* ctx, int enable);
/* 1 = fuse interior step boundaries in rbhip_execute (default), 0 = two-pass */
int rbhip_use_fusion(rbhip_context* ctx, int enable);

/* computeKineticEnergy + getKineticEnergies (RigidBodyKernels.h:80-86,
   ReferenceRigidBodyKernels.cpp:110-121): out2 = {translational, rotational},
   FP64 device reduction */
int rbhip_kinetic_energies(rbhip_context* ctx, double* out2);
/* potential energy of the last force evaluation */
int rbhip_potential_energy(rbhip_context* ctx, double* energy);

/* Refined energies (RigidBodyIntegrator::setComputeRefinedEnergies,
   RigidBodyIntegrator.cpp:34-38): 1 = accumulate, on the last step of every
   rbhip_execute (or each rbhip_part1/rbhip_part2 pair), the CUDA platform's
   COMPMOD terms (platforms/cuda/src/kernels/rigidbodyintegrator.cu:285-292,
   318-320, 380-383).  Enabling allocates 32 B per body and 12 B per free
   atom in single precision (twice that in double); the accumulators start
   empty and every state change (rbhip_set_positions, rbhip_set_velocities)
   empties them again, as the reference's upload zeroes them
   (CudaRigidBodyKernels.cpp:293-335). */
int rbhip_set_refined_energies(rbhip_context* ctx, int enable);
/* getRefinedKineticEnergies (RigidBodyKernels.h:92; CudaRigidBodyKernels.cpp:
   469-477): {translational, rotational} from the position derivative
   estimated over the last step, divided by 6 dt at the current step size;
   {0, 0} before the first step; the plain kinetic energies when refinement
   is off.  Free atoms take the sign of the body terms (the reference's free
   term, CudaRigidBodyKernels.cpp:411-413, counts the backward displacement
   with -1, which gives 4/3 of the kinetic energy at constant velocity). */
int rbhip_refined_kinetic_energies(rbhip_context* ctx, double* out2);
/* getPotentialEnergyRefinement (RigidBodyKernels.h:98; CudaRigidBodyKernels.cpp:
   481-494): -dt^2/24 [sum_free |F|^2/m + sum_bodies |F|^2/M + tau.(I^-1 tau)]
   of the current forces; 0 when refinement is off */
int rbhip_potential_energy_refinement(rbhip_context* ctx, double* energy);

/* Body state of all bodies in reference body order, FP64, 36 values each:
   N, dof, mass, I[3], invI[3], rcm[3], pcm[3], q[4], pi[4], force[3],
   torque[4], twoKt, twoKr, loc (RigidBody.h:18-53).  pcm/pi are the
   post-step values (the folded half-kick is taken out). */
int rbhip_get_body_states(rbhip_context* ctx, double* out);
/* body-fixed atom coordinates d (RigidBodySystem bodyFixedPositions),
   numBodyAtoms x 3, reference body-atom order */
int rbhip_get_body_fixed_positions(rbhip_context* ctx, double* xyz);

/* ---------------------------------------------------------------------- */
/* Launch configuration, autotuning, measurement                           */
/* ---------------------------------------------------------------------- */
/* workgroup size (64..1024, multiple of 64) and grid cap (0 = one workgroup
   per tile) */
int rbhip_set_launch(rbhip_context* ctx, int blockSize, int maxBlocks);
/* lanes per body for the bodies of N atoms (power of two, 1..64) */
int rbhip_set_lanes(rbhip_context* ctx, int atomsPerBody, int lanes);
/* bodies of one shape (mixed precision, on by default): at every build of
   the body frames from positions, the bodies of a size class whose
   body-fixed coordinates agree to within four fp32 ulps of the body's size
   (one molecule type; up to 8 shapes per class, the frame's axis signs) take
   one set of values and read shared rows in the step passes, 12 B per atom
   less HBM traffic per pass.  Takes effect at the next position upload or
   state change.  (A storage choice of this platform: the reference keeps
   every body's coordinates, RigidBodySystem.cpp:94-95.) */
int rbhip_set_shared_frames(rbhip_context* ctx, int enable);
/* times candidate (lanes per body, workgroup size) tiles on a scratch copy of
   the state and keeps the fastest; the context's state is not advanced */
int rbhip_autotune(rbhip_context* ctx, int reps);
/* describes the chosen tiling as text */
int rbhip_describe(rbhip_context* ctx, char* buf, int len);
/* advances the state 2*steps+1 steps on the context's stream: `steps`
   two-pass steps (part1, forces, part2), then the fused path of
   step(steps+1); avgMs[RBHIP_NUM_KERNELS] receives the mean duration of each
   step kernel, from HIP events attached to its own dispatch (launch overhead
   not counted), and totalMs the span of the two-pass steps, gaps included */
int rbhip_time_kernels(rbhip_context* ctx, int steps, double* avgMs, double* totalMs);
/* streams `bytes` through a plain device copy kernel `reps` times (profiler
   calibration of HBM traffic counters); avgMs receives its mean duration */
int rbhip_copy_calibration(size_t bytes, int reps, double* avgMs);

/* ---------------------------------------------------------------------- */
/* Batches: many independent systems stepped by one set of launches        */
/* (no reference counterpart: the reference steps one system per context)  */
/* ---------------------------------------------------------------------- */
int rbhip_batch_create(int device, rbhip_batch** out);
int rbhip_batch_destroy(rbhip_batch* batch);
/* replaces the member list; members must live on the batch's device, have
   the same precision and rotation mode, and hold a built body state */
int rbhip_batch_set_members(rbhip_batch* batch, int count, rbhip_context* const* members);
/* launch stream of the batch (NULL = a library pool stream, the default) */
int rbhip_batch_set_stream(rbhip_batch* batch, rbhip_stream stream);
int rbhip_batch_set_launch(rbhip_batch* batch, int blockSize, int maxBlocks);
int rbhip_batch_use_graph(rbhip_batch* batch, int enable);
int rbhip_batch_use_fusion(rbhip_batch* batch, int enable);
/* steps every member `steps` times: per step one part1, one force and one
   part2 launch over all members' tiles */
int rbhip_batch_execute(rbhip_batch* batch, int steps);
/* per-member {trans, rot} kinetic energies, 2*count doubles */
int rbhip_batch_kinetic_energies(rbhip_batch* batch, double* out);
/* as rbhip_time_kernels, over all members (advances them 2*steps+1 steps) */
int rbhip_batch_time_kernels(rbhip_batch* batch, int steps, double* avgMs, double* totalMs);
/* times candidate launch shapes on the members' real state and keeps the
   fastest: workgroup size, grid cap (0 = one workgroup per tile; a cap makes
   the workgroups loop over the tile list, a persistent grid) and lanes per
   body for each body size present; the members' state is restored */
int rbhip_batch_autotune(rbhip_batch* batch, int reps);
/* describes the batch's launch shape as text */
int rbhip_batch_describe(rbhip_batch* batch, char* buf, int len);

/* ---------------------------------------------------------------------- */
/* Job server: admission by HBM budget and continuous batching             */
/* (no reference counterpart: the reference runs one system per context)   */
/* ---------------------------------------------------------------------- */
typedef struct rbhip_server rbhip_server;

/* device bytes rbhip_create would hold for this system (an upper bound of
   what rbhip_device_bytes reports before bonds are added) and, if numBodies
   is not NULL, the number of rigid bodies */
int rbhip_estimate_bytes(int numAtoms, const int* bodyIndices, int precision, size_t* bytes, int* numBodies);

/* A server on `device`: jobs wait in arrival order and are admitted while the
   device bytes of the resident jobs stay within hbmBudget; resident jobs are
   stepped together as one batch, each round advancing all of them by the
   same count (at most stepsPerRound, never past the first job's end); a job
   that reaches its step count leaves with its final state copied to host
   memory and its device memory freed.  All jobs share the precision and the
   rotation kind (exact or NO-SQUISH) and run with workgroup size blockSize,
   so each job's result equals the job stepped alone with that block size.
   New jobs are set up by four threads of the server and finished jobs
   collected by a fifth, overlapping the queued rounds.
   Errors of the rbhip_server_* calls are described by
   rbhip_server_last_error(). */
const char* rbhip_server_last_error(void);
int rbhip_server_create(int device, size_t hbmBudget, int stepsPerRound, int precision, int blockSize,
                        rbhip_server** out);
int rbhip_server_destroy(rbhip_server* server);
/* the order in which waiting jobs are admitted when the budget has room:
   RBHIP_ADMIT_FIFO (the default) in submission order, or
   RBHIP_ADMIT_LONGEST_FIRST by step count, most steps first (submission
   order among equal counts), which keeps long jobs from running alone at
   the end of a backlog.  Either way admission stops at the first waiting
   job that does not fit.  Applies to the jobs already waiting too. */
enum { RBHIP_ADMIT_FIFO = 0, RBHIP_ADMIT_LONGEST_FIRST = 1 };
int rbhip_server_set_admission_order(rbhip_server* server, int order);
/* queues a job; velocities and the tether may be NULL.  The arrays are not
   copied: the server reads them when it sets the job up, so they must stay
   valid and unchanged until rbhip_server_take has returned the job (or the
   server is destroyed).  Rejects a job larger than the whole budget. */
int rbhip_server_submit(rbhip_server* server, int numAtoms, const double* masses, const int* bodyIndices,
                        const double* positions, const double* velocities, const double* tetherK,
                        const double* tetherR0, double stepSize, int rotationMode, long long steps,
                        long long* jobId);
/* one round: hands up to 8 waiting jobs (in arrival order, within the
   budget) to the admission thread, hands the jobs that reached their step
   count to the collection thread, and queues the next batched advance of
   the resident and newly set-up jobs without waiting for it (it waits only
   to keep at most two rounds queued, or, with nothing resident, for the
   first admission); *finished = jobs completed since the previous call.
   Jobs complete on the collection thread, so rbhip_server_completed can
   list more of them between calls. */
int rbhip_server_run_round(rbhip_server* server, int* finished);
/* counts3 = {waiting, resident (including finished jobs not yet collected,
   which still hold device memory), completed and not yet taken}; device
   bytes of the resident jobs; body-steps done so far.  Any output may be
   NULL. */
int rbhip_server_status(rbhip_server* server, long long* counts3, size_t* bytesInUse, long long* bodySteps);
/* rounds2 = {rounds queued so far, those cut below steps-per-round by a job
   reaching its step count}; seconds6 = {admitting (summed over the admission
   threads), collecting finished jobs (collection thread, includes waiting
   for the device), waiting for the oldest queued round, updating the batch
   members, queueing the steps (these three on the run_round thread), the
   retired rounds' duration on the device}.  Either output may be NULL. */
int rbhip_server_profile(rbhip_server* server, long long* rounds2, double* seconds6);
/* ids of completed jobs not yet taken, up to max */
int rbhip_server_completed(rbhip_server* server, long long* ids, int max, int* count);
/* final positions and velocities (numAtoms x 3, original order), {KE_t, KE_r}
   and times3 = {seconds waiting for admission, seconds resident, completion
   latency from submission} of a completed job, which is then released; any
   output may be NULL */
int rbhip_server_take(rbhip_server* server, long long jobId, double* positions, double* velocities, double* ke2,
                      double* times3);

#ifdef __cplusplus
}
#endif
#endif
