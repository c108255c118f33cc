/*
 * rigidbody_hip_plugin.h — C entry points of the host plugin library
 * (librigidbody_hip.so): plugin registration, and a handle-based C view of
 * System + RigidBodyIntegrator + Context on the HIP platform for bindings.
 *
 * Conventions as rbhip.h: int status (0 = OK, negative on failure),
 * rbplugin_last_error() describes the last failure on the calling thread;
 * host arrays of vectors are xyz-interleaved FP64 in original atom order.
 */
#ifndef RIGIDBODY_HIP_PLUGIN_H_
#define RIGIDBODY_HIP_PLUGIN_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Plugin registration.  registerPlatforms / registerKernelFactories are the
   symbols a plugin loader calls (reference:
   platforms/reference/src/ReferenceRigidBodyKernelFactory.cpp:41-52,
   platforms/cuda/src/CudaRigidBodyKernelFactory.cpp:41-55); the HIP platform
   is this library's own, so registerPlatforms registers it.
   registerRigidBodyHipKernelFactories registers the platform if missing and
   then its kernel factory, for direct use without a loader (as
   registerRigidBodyCudaKernelFactories, CudaRigidBodyKernelFactory.cpp:57-65). */
void registerPlatforms(void);
void registerKernelFactories(void);
void registerRigidBodyHipKernelFactories(void);

typedef struct rbplugin_sim rbplugin_sim;

const char* rbplugin_last_error(void);

/* System(masses) + forces + RigidBodyIntegrator(stepSize, bodyIndices) with
   setRotationMode(rotationMode) + Context on the HIP platform (device,
   precision 0 = mixed, 1 = double).  Forces: numBonds HarmonicBondForce bonds
   (pairs, length, k; may be 0) and, if tetherK != NULL, a TetherForce on
   every atom with tetherK[i] != 0 (anchors tetherR0, xyz).
   RigidBodyIntegrator.cpp:18-51 + the Context constructor. */
int rbplugin_create(int numAtoms, const double* masses, const int* bodyIndices, double stepSize,
                    int rotationMode, int precision, int device, int numBonds, const int* bondPairs,
                    const double* bondLength, const double* bondK, const double* tetherK,
                    const double* tetherR0, rbplugin_sim** out);
/* as rbplugin_create, with RigidBodyIntegrator::setComputeRefinedEnergies
   (RigidBodyIntegrator.cpp:34-38) called before the Context binds it */
int rbplugin_create_ex(int numAtoms, const double* masses, const int* bodyIndices, double stepSize,
                       int rotationMode, int computeRefinedEnergies, int precision, int device, int numBonds,
                       const int* bondPairs, const double* bondLength, const double* bondK, const double* tetherK,
                       const double* tetherR0, rbplugin_sim** out);
/* A force field evaluated on the host (rbhost::HostForce): at each force
   evaluation the callback receives the positions of all atoms (xyz, FP64,
   original order), adds its forces to `forces` (zero on entry) and returns
   its potential energy.  It is called between part 1 and part 2 of every
   step (RigidBodyIntegrator::step then runs one execute() per step,
   ReferenceRigidBodyKernels.cpp:98-104) and at state changes. */
typedef double (*rbplugin_force_fn)(void* user, int numAtoms, const double* positions, double* forces);
/* as rbplugin_create_ex, with hostForce (may be NULL) added to the System */
int rbplugin_create_with_force(int numAtoms, const double* masses, const int* bodyIndices, double stepSize,
                               int rotationMode, int computeRefinedEnergies, int precision, int device,
                               int numBonds, const int* bondPairs, const double* bondLength, const double* bondK,
                               const double* tetherK, const double* tetherR0, rbplugin_force_fn hostForce,
                               void* user, rbplugin_sim** out);
int rbplugin_destroy(rbplugin_sim* sim);
/* Context::setPositions / setVelocities (-> RigidBodyIntegrator::stateChanged) */
int rbplugin_set_positions(rbplugin_sim* sim, const double* xyz);
int rbplugin_set_velocities(rbplugin_sim* sim, const double* xyz);
/* RigidBodyIntegrator::step(steps) */
int rbplugin_step(rbplugin_sim* sim, int steps);
/* Integrator::setStepSize */
int rbplugin_set_step_size(rbplugin_sim* sim, double stepSize);
/* Context::getState(Positions | Velocities | Forces | Energy): any output may
   be NULL; energies3 = {kinetic, potential, time} */
int rbplugin_get_state(rbplugin_sim* sim, double* positions, double* velocities, double* forces,
                       double* energies3);
/* RigidBodyIntegrator::getKineticEnergies, getRefinedKineticEnergies,
   getPotentialEnergyRefinement: out5 = {KEt, KEr, refined KEt, refined KEr,
   PE refinement} */
int rbplugin_energies(rbplugin_sim* sim, double* out5);
/* RigidBodySystem counts: numFree, numBodies, numActualAtoms, numBodyAtoms,
   numDOF (RigidBodySystem.h:36-40) */
int rbplugin_counts(rbplugin_sim* sim, int* out5);
/* RigidBodySystem::getAtomIndex for i in [0, numActualAtoms) */
int rbplugin_atom_index(rbplugin_sim* sim, int* out);
/* RigidBodySystem::getRigidBody(i): N, dof, mass, I[3], invI[3], rcm[3],
   pcm[3], q[4], pi[4], force[3], torque[4], twoKt, twoKr, loc (33 values) */
int rbplugin_rigid_body(rbplugin_sim* sim, int index, double* out33);

#ifdef __cplusplus
}
#endif
#endif
