// The kernel interface of the rigid-body step, as
// openmmapi/include/RigidBodyKernels.h:47-99 (name, the seven virtuals and
// their argument meaning), on the host API of HostApi.h.
#pragma once

#include <string>
#include <vector>

#include "HostApi.h"
#include "RigidBodyIntegrator.h"

namespace RigidBodyPlugin {

class IntegrateRigidBodyStepKernel : public rbhost::KernelImpl {
public:
    static std::string Name() { return "IntegrateRigidBodyStep"; }
    IntegrateRigidBodyStepKernel(std::string name, const rbhost::Platform& platform)
        : rbhost::KernelImpl(name, platform) {}
    virtual void initialize(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) = 0;
    virtual void uploadBodySystem(RigidBodySystem& bodySystem) = 0;
    // one time step
    virtual void execute(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) = 0;
    // `steps` time steps.  The reference integrator calls execute() once per
    // step (RigidBodyIntegrator.cpp:99-100), which this default keeps; a
    // platform may advance all of them in one call.
    virtual void execute(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator, int steps) {
        for (int i = 0; i < steps; i++) execute(context, integrator);
    }
    virtual double computeKineticEnergy(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) = 0;
    virtual std::vector<double> getKineticEnergies(const RigidBodyIntegrator& integrator) = 0;
    virtual std::vector<double> getRefinedKineticEnergies(const RigidBodyIntegrator& integrator) = 0;
    virtual double getPotentialEnergyRefinement(const RigidBodyIntegrator& integrator) = 0;
};

}  // namespace RigidBodyPlugin
