// The kernel interface of the rigid-body step, as
// openmmapi/include/RigidBodyKernels.h:47-99 (name, the seven virtuals and
// their argument meaning), on the host API of HostApi.h.
//
// MultiStepKernel is not part of that interface: a platform kernel may also
// implement it to advance several steps in one call when the context's
// forces are its own, and RigidBodyIntegrator::step then uses it in place of
// one execute() per step.
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
    // one time step: part 1, the context's force evaluation, part 2
    virtual void execute(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) = 0;
    virtual double computeKineticEnergy(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) = 0;
    virtual std::vector<double> getKineticEnergies(const RigidBodyIntegrator& integrator) = 0;
    virtual std::vector<double> getRefinedKineticEnergies(const RigidBodyIntegrator& integrator) = 0;
    virtual double getPotentialEnergyRefinement(const RigidBodyIntegrator& integrator) = 0;
};

class MultiStepKernel {
public:
    virtual ~MultiStepKernel() = default;
    // true if `steps` calls of execute() can run as one executeSteps() call:
    // every force of the context is evaluated by the platform kernel's own
    // device code, so no host force evaluation has to run between steps
    virtual bool canExecuteSteps(rbhost::ContextImpl& context) const = 0;
    // the same trajectory as `steps` calls of execute()
    virtual void executeSteps(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator, int steps) = 0;
};

}  // namespace RigidBodyPlugin
