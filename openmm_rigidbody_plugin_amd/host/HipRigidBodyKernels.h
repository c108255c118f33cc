// HipIntegrateRigidBodyStepKernel: the rigid-body step on a gfx950 device
// through the rbhip C ABI.  Counterpart of the reference's
// ReferenceIntegrateRigidBodyStepKernel (platforms/reference/include/
// ReferenceRigidBodyKernels.h) and CudaIntegrateRigidBodyStepKernel
// (platforms/cuda/include/CudaRigidBodyKernels.h).
#pragma once

#include "HipPlatform.h"
#include "RigidBodyKernels.h"

namespace RigidBodyPlugin {

class HipIntegrateRigidBodyStepKernel : public IntegrateRigidBodyStepKernel, public MultiStepKernel {
public:
    HipIntegrateRigidBodyStepKernel(std::string name, const rbhost::Platform& platform,
                                    rbhost::HipPlatform::PlatformData& data)
        : IntegrateRigidBodyStepKernel(name, platform), data(data) {}
    // creates the device context: bookkeeping, internal atom order, device
    // arrays and the force field of the System
    void initialize(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) override;
    // the body state is built on the device when positions or velocities are
    // set; this binds the host view to it
    void uploadBodySystem(RigidBodySystem& bodySystem) override;
    // part 1 on the device, the context's force evaluation, part 2 on the
    // device (ReferenceRigidBodyKernels.cpp:82-108)
    void execute(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) override;
    // all forces on the device (no HostForce): the steps run in one call,
    // with fused step boundaries replayed from HIP graphs
    bool canExecuteSteps(rbhost::ContextImpl& context) const override;
    void executeSteps(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator, int steps) override;
    double computeKineticEnergy(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) override;
    std::vector<double> getKineticEnergies(const RigidBodyIntegrator& integrator) override;
    // with setComputeRefinedEnergies(true), the CUDA platform's refined values
    // (CudaRigidBodyKernels.cpp:469-494), otherwise the plain kinetic energies
    // and 0, as the reference platform returns (ReferenceRigidBodyKernels.cpp:123-132)
    std::vector<double> getRefinedKineticEnergies(const RigidBodyIntegrator& integrator) override;
    double getPotentialEnergyRefinement(const RigidBodyIntegrator& integrator) override;

private:
    void syncStepSize(const RigidBodyIntegrator& integrator);
    // the reference reads the tolerance at every step (ReferenceRigidBodyKernels.cpp:85)
    void syncTolerance(const RigidBodyIntegrator& integrator);
    rbhost::HipPlatform::PlatformData& data;
    double stepSize = 0.0;
    double tolerance = -1.0;
};

class HipRigidBodyKernelFactory : public rbhost::KernelFactory {
public:
    rbhost::KernelImpl* createKernelImpl(std::string name, const rbhost::Platform& platform,
                                         rbhost::ContextImpl& context) const override;
};

}  // namespace RigidBodyPlugin
