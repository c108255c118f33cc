// RigidBodyIntegrator: symplectic, time-reversible integration of rigid bodies
// and free atoms, with exact (Jacobi elliptic) or NO-SQUISH rotations.
//
// Public interface as openmmapi/include/RigidBodyIntegrator.h:28-134 (same
// names, argument meaning and exceptions), on the host API of HostApi.h.
#pragma once

#include <string>
#include <vector>

#include "HostApi.h"
#include "RigidBodySystem.h"

namespace RigidBodyPlugin {

class RigidBodyIntegrator : public rbhost::Integrator {
public:
    // stepSize in ps; bodyIndices[i] = rigid body of atom i (<= 0: free atom)
    explicit RigidBodyIntegrator(double stepSize, const std::vector<int>& bodyIndices);
    // 0 = exact rotation (default); n > 0 = NO-SQUISH with n substeps
    void setRotationMode(int mode);
    int getRotationMode() const { return rotationMode; }
    void setComputeRefinedEnergies(bool compute);
    bool getComputeRefinedEnergies() const { return computeRefinedEnergies; }
    // advances `steps` time steps (RigidBodyIntegrator.cpp:96-101)
    void step(int steps) override;
    std::vector<int> getBodyIndices() const { return bodyIndices; }
    const RigidBodySystem& getRigidBodySystem() const { return bodySystem; }
    // {translational, rotational} kinetic energy
    std::vector<double> getKineticEnergies();
    std::vector<double> getRefinedKineticEnergies();
    double getPotentialEnergyRefinement();

protected:
    void initialize(rbhost::ContextImpl& context) override;
    void cleanup() override;
    std::vector<std::string> getKernelNames() override;
    void stateChanged(rbhost::State::DataType changed) override;
    double computeKineticEnergy() override;

private:
    std::vector<int> bodyIndices;
    RigidBodySystem bodySystem;
    rbhost::Kernel kernel;
    int rotationMode;
    bool computeRefinedEnergies;
};

}  // namespace RigidBodyPlugin
