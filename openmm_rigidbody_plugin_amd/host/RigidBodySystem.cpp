// Host-side view of the system of rigid bodies (see RigidBodySystem.h).
#include "RigidBodySystem.h"

#include <algorithm>

#include "HipPlatform.h"

namespace RigidBodyPlugin {

using rbhost::Exception;
using rbhost::Vec3;

std::vector<int> cleanBodyIndices(const std::vector<int>& bodyIndices) {
    std::vector<int> labels;
    for (int b : bodyIndices)
        if (b > 0) labels.push_back(b);
    std::sort(labels.begin(), labels.end());
    labels.erase(std::unique(labels.begin(), labels.end()), labels.end());
    std::vector<int> out(bodyIndices.size(), 0);
    for (size_t i = 0; i < bodyIndices.size(); i++)
        if (bodyIndices[i] > 0)
            out[i] = (int)(std::lower_bound(labels.begin(), labels.end(), bodyIndices[i]) - labels.begin()) + 1;
    return out;
}

void RigidBodySystem::initialize(rbhost::ContextImpl& context, const std::vector<int>& bodyIndices, int mode) {
    const rbhost::System& system = context.getSystem();
    const int numAtoms = system.getNumParticles();
    bodyIndex = cleanBodyIndices(bodyIndices);
    rotationMode = mode;
    numBodies = 0;
    for (int b : bodyIndex) numBodies = std::max(numBodies, b);
    // RigidBodySystem.cpp:64-79: virtual sites are not actual atoms, and
    // neither they nor massless particles are free or body atoms
    numActualAtoms = numAtoms;
    for (int i = 0; i < numAtoms; i++)
        if (system.isVirtualSite(i)) numActualAtoms--;
    atomIndex.assign(numActualAtoms, 0);
    bodyN.assign(numBodies, 0);
    numFree = 0;
    for (int i = 0; i < numAtoms; i++) {
        if (system.isVirtualSite(i) || system.getParticleMass(i) == 0.0) continue;
        if (bodyIndex[i] == 0) atomIndex[numFree++] = i;
        else bodyN[bodyIndex[i] - 1]++;
    }
    numBodyAtoms = numActualAtoms - numFree;
    bodyLoc.assign(numBodies, 0);
    int loc = 0;
    for (int b = 0; b < numBodies; b++) {
        bodyLoc[b] = loc;
        loc += bodyN[b];
    }
    std::vector<int> fill(numBodies, 0);
    for (int i = 0; i < numAtoms; i++) {
        int b = bodyIndex[i];
        if (b > 0) atomIndex[numFree + bodyLoc[b - 1] + fill[b - 1]++] = i;
    }
    numConstraints = system.getNumConstraints();
    for (int c = 0; c < numConstraints; c++) {
        int a1, a2;
        double dist;
        system.getConstraintParameters(c, a1, a2, dist);
        if (bodyIndex.at(a1) != 0 || bodyIndex.at(a2) != 0)
            throw Exception("Constraints involving rigid-body atoms are not allowed");
    }
    numDOF = 0;
    device = nullptr;
}

rbhip_context* RigidBodySystem::require() const {
    if (!device) throw Exception("The rigid-body system has no device state yet");
    return device;
}

void RigidBodySystem::update(rbhost::ContextImpl&, bool geometry, bool) {
    if (!geometry) return;
    int counts[5];
    rbhost::rbhip_check(rbhip_get_counts(require(), counts));
    // numDOF = numFree - numConstraints + sum of body dof (RigidBodySystem.cpp:130-134),
    // as the device context counts it
    numDOF = counts[4];
}

Vec3 RigidBodySystem::getBodyFixedPosition(int i) const {
    if (i < 0 || i >= numBodyAtoms) throw Exception("Body atom index out of range");
    std::vector<double> d(3 * (size_t)numBodyAtoms);
    rbhost::rbhip_check(rbhip_get_body_fixed_positions(require(), d.data()));
    return Vec3(d[3 * (size_t)i], d[3 * (size_t)i + 1], d[3 * (size_t)i + 2]);
}

RigidBody RigidBodySystem::getRigidBody(int i) const {
    if (i < 0 || i >= numBodies) throw Exception("Body index out of range");
    std::vector<double> s(36 * (size_t)numBodies);
    rbhost::rbhip_check(rbhip_get_body_states(require(), s.data()));
    const double* o = s.data() + 36 * (size_t)i;
    RigidBody b;
    b.N = (int)o[0];
    b.dof = (int)o[1];
    b.mass = o[2];
    b.I = Vec3(o[3], o[4], o[5]);
    b.invI = Vec3(o[6], o[7], o[8]);
    b.rcm = Vec3(o[9], o[10], o[11]);
    b.pcm = Vec3(o[12], o[13], o[14]);
    for (int c = 0; c < 4; c++) b.q[c] = o[15 + c];
    for (int c = 0; c < 4; c++) b.pi[c] = o[19 + c];
    b.force = Vec3(o[23], o[24], o[25]);
    for (int c = 0; c < 4; c++) b.torque[c] = o[26 + c];
    b.twoKt = o[30];
    b.twoKr = o[31];
    b.loc = bodyLoc[i];
    b.invMass = 1.0 / b.mass;
    return b;
}

double RigidBodySystem::getTranslationalEnergy() const {
    double ke[2];
    rbhost::rbhip_check(rbhip_kinetic_energies(require(), ke));
    return ke[0];
}

double RigidBodySystem::getRotationalEnergy() const {
    double ke[2];
    rbhost::rbhip_check(rbhip_kinetic_energies(require(), ke));
    return ke[1];
}

}  // namespace RigidBodyPlugin
