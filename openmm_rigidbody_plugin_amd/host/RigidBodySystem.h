// Host-side view of the system of rigid bodies and free atoms.
//
// Mirrors RigidBodyPlugin::RigidBodySystem (openmmapi/include/RigidBodySystem.h:27-61):
// the integer bookkeeping (cleanBodyIndices, atomIndex, counts, DOF) is built
// on the host with the reference's rules, and the per-body dynamic state is
// read from the device library on request (the device builds and advances
// it; the host copy never goes stale).
#pragma once

#include <vector>

#include "HostApi.h"
#include "rbhip.h"

namespace RigidBodyPlugin {

// one body's state (openmmapi/include/RigidBody.h:18-53), FP64
struct RigidBody {
    int N = 0;                  // number of atoms
    int dof = 0;                // degrees of freedom (5 linear, 6 otherwise)
    double mass = 0.0;
    rbhost::Vec3 I;             // principal moments of inertia (descending)
    rbhost::Vec3 rcm;           // centre-of-mass position
    rbhost::Vec3 pcm;           // centre-of-mass momentum
    double q[4] = {1, 0, 0, 0}; // unit orientation quaternion
    double pi[4] = {0, 0, 0, 0};// quaternion-conjugated momentum, pi = 2 B(q) L
    rbhost::Vec3 force;         // resultant force
    double torque[4] = {0, 0, 0, 0}; // quaternion-frame torque C(q) tau
    int loc = 0;                // location of the first atom in atomIndex (after the free atoms)
    double invMass = 0.0;
    rbhost::Vec3 invI;
    double twoKt = 0.0, twoKr = 0.0;
};

class RigidBodySystem {
public:
    // RigidBodySystem::initialize (openmmapi/src/RigidBodySystem.cpp:55-114)
    void initialize(rbhost::ContextImpl& context, const std::vector<int>& bodyIndices, int rotationMode);
    // RigidBodySystem::update (RigidBodySystem.cpp:120-142): the device has
    // rebuilt the body state; refresh the DOF count from it
    void update(rbhost::ContextImpl& context, bool geometry, bool velocities);
    // the device context holding the body state (set by the HIP kernel)
    void attach(rbhip_context* device) { this->device = device; }

    int getNumDOF() const { return numDOF; }
    int getNumFree() const { return numFree; }
    int getNumBodies() const { return numBodies; }
    int getNumActualAtoms() const { return numActualAtoms; }
    int getNumBodyAtoms() const { return numBodyAtoms; }
    int getAtomIndex(int i) const { return atomIndex.at(i); }
    int getBodyIndex(int atom) const { return bodyIndex.at(atom); }
    int getRotationMode() const { return rotationMode; }
    // read from the device: body-fixed coordinates of body atom i (reference
    // body-atom order) and the state of body i (reference body order)
    rbhost::Vec3 getBodyFixedPosition(int i) const;
    RigidBody getRigidBody(int i) const;
    double getTranslationalEnergy() const;
    double getRotationalEnergy() const;
    double getKineticEnergy() const { return getTranslationalEnergy() + getRotationalEnergy(); }

private:
    rbhip_context* require() const;
    int rotationMode = 0;
    std::vector<int> bodyIndex, atomIndex, bodyN, bodyLoc;
    int numFree = 0, numBodies = 0, numActualAtoms = 0, numBodyAtoms = 0, numDOF = 0;
    int numConstraints = 0;
    rbhip_context* device = nullptr;
};

// cleanBodyIndices (RigidBodySystem.cpp:28-49): distinct positive labels become
// 1..B in ascending label order, labels <= 0 become 0 (free atoms)
std::vector<int> cleanBodyIndices(const std::vector<int>& bodyIndices);

}  // namespace RigidBodyPlugin
