// HipIntegrateRigidBodyStepKernel (see HipRigidBodyKernels.h).
#include "HipRigidBodyKernels.h"

namespace RigidBodyPlugin {

using rbhost::Exception;
using rbhost::rbhip_check;

void HipIntegrateRigidBodyStepKernel::initialize(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) {
    stepSize = integrator.getStepSize();
    data.createSimulation(context.getSystem(), integrator.getBodyIndices(), stepSize, integrator.getRotationMode());
    tolerance = -1.0;
    syncTolerance(integrator);
    if (integrator.getComputeRefinedEnergies()) rbhip_check(rbhip_set_refined_energies(data.require(), 1));
}

void HipIntegrateRigidBodyStepKernel::uploadBodySystem(RigidBodySystem& bodySystem) {
    bodySystem.attach(data.require());
}

void HipIntegrateRigidBodyStepKernel::syncStepSize(const RigidBodyIntegrator& integrator) {
    double dt = integrator.getStepSize();
    if (dt != stepSize) {
        rbhip_check(rbhip_set_step_size(data.require(), dt));
        stepSize = dt;
    }
}

void HipIntegrateRigidBodyStepKernel::execute(rbhost::ContextImpl& context, const RigidBodyIntegrator& integrator) {
    syncStepSize(integrator);
    syncTolerance(integrator);
    rbhip_context* sim = data.require();
    rbhip_check(rbhip_part1(sim));
    context.calcForcesAndEnergy(true, false);
    rbhip_check(rbhip_part2(sim));
}

void HipIntegrateRigidBodyStepKernel::syncTolerance(const RigidBodyIntegrator& integrator) {
    double tol = integrator.getConstraintTolerance();
    if (tol != tolerance) {
        rbhip_check(rbhip_set_constraint_tolerance(data.require(), tol));
        tolerance = tol;
    }
}

bool HipIntegrateRigidBodyStepKernel::canExecuteSteps(rbhost::ContextImpl& context) const {
    return !data.hasHostForces(context.getSystem());
}

void HipIntegrateRigidBodyStepKernel::executeSteps(rbhost::ContextImpl&, const RigidBodyIntegrator& integrator,
                                                   int steps) {
    syncStepSize(integrator);
    syncTolerance(integrator);
    rbhip_check(rbhip_execute(data.require(), steps));
}

double HipIntegrateRigidBodyStepKernel::computeKineticEnergy(rbhost::ContextImpl&, const RigidBodyIntegrator&) {
    double ke[2];
    rbhip_check(rbhip_kinetic_energies(data.require(), ke));
    return ke[0] + ke[1];
}

std::vector<double> HipIntegrateRigidBodyStepKernel::getKineticEnergies(const RigidBodyIntegrator&) {
    double ke[2];
    rbhip_check(rbhip_kinetic_energies(data.require(), ke));
    return {ke[0], ke[1]};
}

// both divide by the integrator's current step size, as the reference does
// (CudaRigidBodyKernels.cpp:159, 492)
std::vector<double> HipIntegrateRigidBodyStepKernel::getRefinedKineticEnergies(const RigidBodyIntegrator& integrator) {
    syncStepSize(integrator);
    double ke[2];
    rbhip_check(rbhip_refined_kinetic_energies(data.require(), ke));
    return {ke[0], ke[1]};
}

double HipIntegrateRigidBodyStepKernel::getPotentialEnergyRefinement(const RigidBodyIntegrator& integrator) {
    syncStepSize(integrator);
    double u = 0.0;
    rbhip_check(rbhip_potential_energy_refinement(data.require(), &u));
    return u;
}

rbhost::KernelImpl* HipRigidBodyKernelFactory::createKernelImpl(std::string name, const rbhost::Platform& platform,
                                                                rbhost::ContextImpl& context) const {
    auto& data = rbhost::HipPlatform::data(context);
    if (name == IntegrateRigidBodyStepKernel::Name()) return new HipIntegrateRigidBodyStepKernel(name, platform, data);
    throw Exception("Tried to create kernel with illegal kernel name '" + name + "'");
}

}  // namespace RigidBodyPlugin
