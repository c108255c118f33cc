/*
 * rbref — CPU FP64 restatement of the reference RigidBodyIntegrator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (openmm_rigidbody_plugin_amd) never links or calls
 * it.
 *
 * What it restates (reference files, read-only, cited per function in
 * rbref.cpp):
 *   openmmapi/src/RigidBodyIntegrator.cpp   initialize / stateChanged / step
 *   openmmapi/src/RigidBodySystem.cpp       cleanBodyIndices, initialize,
 *                                           update, integratePart1/2, KE
 *   openmmapi/src/RigidBody.cpp             geometry/dynamics build, rotations
 *   openmmapi/src/MatVec.cpp                Quat B/C/Bt/Ct/B1..B3/A/At, Shepperd
 *   openmmapi/src/eigenDecomposition.cpp    Smith eigenvalues, Kopp eigenvectors
 *   openmmapi/include/internal/ellipticFunctions.h  (re-derived from Carlson
 *                                           1995 and DLMF 22.20, see rbref.cpp)
 *   platforms/reference/src/ReferenceRigidBodyKernels.cpp  execute()
 *
 * Pinning status of each function is recorded in the header comment of
 * rbref.cpp and in the tests that check it.
 *
 * Deviation flags (RBREF_FIX_*): with flags == 0 the restatement reproduces the
 * reference's documented defects; RBREF_FIX_ALL is the policy the HIP path
 * implements.
 */
#ifndef RBREF_H_
#define RBREF_H_

#ifdef __cplusplus
extern "C" {
#endif

/* zero the centre-of-mass momentum before buildDynamics accumulates it
   (reference RigidBody.cpp:126-130 never resets it) */
#define RBREF_FIX_PCM_RESET      1u
/* linear bodies: body frame rows (v, u x v, u), so d = A delta lies on e_z
   (reference RigidBody.cpp:96 transposes it) */
#define RBREF_FIX_LINEAR_FRAME   2u
/* exact mode, dof 5 bodies: closed-form linear-rotor rotation instead of the
   elliptic solution that divides by I0-I1 = 0 (RigidBody.cpp:254-255) */
#define RBREF_FIX_LINEAR_EXACT   4u
/* velocity-only rebuilds use the live body geometry instead of the
   integrator's stale copy (RigidBodyIntegrator.cpp:70-72 + TODO.md:5) */
#define RBREF_FIX_STALE_SYSTEM   8u
/* linear bodies: the axis-orthogonal frame vector starts from the axis of
   the smallest |u_k| (reference RigidBody.cpp:51-58 takes the smallest signed
   component and returns NaN for a body axis along -e_k) */
#define RBREF_FIX_ORTHONORMAL    16u
/* refined kinetic energy of free atoms: the backward displacement enters the
   position derivative with +1, as in the body terms (the CUDA platform's
   CudaRigidBodyKernels.cpp:411-413 uses -1, which gives 4/3 of the kinetic
   energy at constant velocity) */
#define RBREF_FIX_REFINED_FREE   32u
#define RBREF_FIX_ALL            63u
/* planted mistake, for tests that show a check has teeth: the body-atom
   velocities are rebuilt as delta_j x w_s instead of w_s x delta_j
   (RigidBody.cpp:166); never part of a policy */
#define RBREF_PLANT_VELOCITY_SIGN 1024u

typedef struct rbref_sim rbref_sim;

const char* rbref_last_error(void);

/* ---- system / integrator (RigidBodyIntegrator + Reference platform) ---- */
/* Builds System(masses) + RigidBodyIntegrator(dt, bodyIndices) +
   setRotationMode(mode) + Context.  Returns NULL and sets the error on the
   exceptions RigidBodyIntegrator.cpp:27-47 / RigidBodySystem.cpp:111 raise. */
rbref_sim* rbref_create(int numAtoms, const double* masses, const int* bodyIndices,
                        double dt, int rotationMode, unsigned flags);
void rbref_destroy(rbref_sim* s);
/* force field: per-atom harmonic tether to r0 (k per atom) and harmonic bonds */
int rbref_set_tether(rbref_sim* s, const double* k, const double* r0);
int rbref_add_bond(rbref_sim* s, int a, int b, double length, double k);
/* Context::setPositions / setVelocities → stateChanged (xyz interleaved, original order) */
/* a distance constraint between two free atoms (System::addConstraint);
   refused for body atoms, as RigidBodySystem.cpp:107-112 does.  Takes effect
   in the next step; the DOF count follows at the next state update. */
int rbref_add_constraint(rbref_sim* s, int a, int b, double distance);
int rbref_set_constraint_tolerance(rbref_sim* s, double tol);
/* an OpenMM virtual site on a massless particle: positions after part 1 and
   the constraints, forces moved to the parents after each force evaluation.
   weights: {w1, w2} (two-particle average), {w1, w2, w3} (three-particle
   average), {w12, w13, wCross} (out of plane), or 12 values for local
   coordinates: origin weights, x weights, y weights (3 each) and the local
   position */
#define RBREF_SITE_TWO_AVERAGE       1
#define RBREF_SITE_THREE_AVERAGE     2
#define RBREF_SITE_OUT_OF_PLANE      3
#define RBREF_SITE_LOCAL_COORDINATES 4
int rbref_add_virtual_site(rbref_sim* s, int type, int site, const int* parents, const double* weights);
long rbref_constraint_failures(rbref_sim* s);   /* clusters not converged, so far */
int rbref_set_positions(rbref_sim* s, const double* xyz);
int rbref_set_velocities(rbref_sim* s, const double* xyz);
int rbref_set_step_size(rbref_sim* s, double dt);
/* RigidBodyIntegrator::step */
int rbref_step(rbref_sim* s, int steps);
int rbref_get_positions(rbref_sim* s, double* xyz);
int rbref_get_velocities(rbref_sim* s, double* xyz);
int rbref_get_forces(rbref_sim* s, double* xyz);
/* computeKineticEnergy() then getKineticEnergies(): out = {trans, rot} */
int rbref_kinetic_energies(rbref_sim* s, double* out2);
double rbref_potential_energy(rbref_sim* s);
/* refined energies (setComputeRefinedEnergies): the CUDA platform's COMPMOD
   accumulators, zeroed on enable and on every state change */
int rbref_set_refined(rbref_sim* s, int enable);
int rbref_refined_kinetic_energies(rbref_sim* s, double* out2);   /* plain KE when off */
double rbref_potential_energy_refinement(rbref_sim* s);           /* 0 when off */
double rbref_time(rbref_sim* s);
/* counts: numFree, numBodies, numActualAtoms, numBodyAtoms, numDOF */
int rbref_counts(rbref_sim* s, int* out5);
int rbref_atom_index(rbref_sim* s, int* out /* numActualAtoms */);
int rbref_body_index(rbref_sim* s, int* out /* numAtoms */);
/* body state of body k (0-based, reference numbering), 36 doubles:
   N, dof, mass, I[3], invI[3], rcm[3], pcm[3], q[4], pi[4], force[3],
   torque[4], twoKt, twoKr, loc */
int rbref_body_state(rbref_sim* s, int k, double* out36);
/* body-fixed coordinates of all body atoms (numBodyAtoms x 3) */
int rbref_body_fixed_positions(rbref_sim* s, double* xyz);

/* ---- unit-level restated functions (for golden-vector checks) ---- */
double rbref_carlson_rf(double x, double y, double z);
double rbref_carlson_rj(double x, double y, double z, double p);
double rbref_carlson_rc(double x, double y);
double rbref_omega(double x, double n, double m);
void   rbref_jacobi(double u, double m, double* sn, double* cn, double* dn);
/* Smith eigenvalues (descending) + Kopp eigenvectors (rows) of symmetric A */
void   rbref_eigen(const double* A9, double* w3, double* V9);
void   rbref_quat_from_matrix(const double* A9, double* q4);
/* A(q) x (space -> body) and A(q)^T x (body -> space) */
void   rbref_quat_rotate(const double* q4, const double* x3, int transpose, double* y3);
void   rbref_clean_body_indices(int n, const int* in, int* out);
/* free rotation of (q, pi) over dt: mode 0 exact, n NO-SQUISH(n) */
void   rbref_rotate(double* q4, double* pi4, const double* I3, const double* invI3,
                    int dof, double dt, int mode, unsigned flags);

#ifdef __cplusplus
}
#endif
#endif
