// RigidBodyIntegrator (openmmapi/src/RigidBodyIntegrator.cpp:18-101 semantics).
#include "RigidBodyIntegrator.h"

#include "RigidBodyKernels.h"

namespace RigidBodyPlugin {

using rbhost::Exception;

RigidBodyIntegrator::RigidBodyIntegrator(double stepSize, const std::vector<int>& bodyIndices)
    : bodyIndices(bodyIndices), rotationMode(0), computeRefinedEnergies(false) {
    setStepSize(stepSize);
    setConstraintTolerance(1e-5);
}

void RigidBodyIntegrator::setRotationMode(int mode) {
    if (mode < 0) throw Exception("Rotation mode cannot be negative");
    if (owner != nullptr) throw Exception("Cannot set rotation mode: integrator already bound to a context");
    rotationMode = mode;
}

void RigidBodyIntegrator::setComputeRefinedEnergies(bool compute) {
    if (owner != nullptr)
        throw Exception("Cannot set refined energy computation: integrator already bound to a context");
    computeRefinedEnergies = compute;
}

void RigidBodyIntegrator::initialize(rbhost::ContextImpl& contextRef) {
    if (owner != nullptr && &contextRef.getOwner() != owner)
        throw Exception("This Integrator is already bound to a context");
    if ((size_t)contextRef.getSystem().getNumParticles() != bodyIndices.size())
        throw Exception("Number of body indices differs from that of atoms in Context");
    context = &contextRef;
    owner = &contextRef.getOwner();
    bodySystem.initialize(contextRef, bodyIndices, rotationMode);
    kernel = contextRef.getPlatform().createKernel(IntegrateRigidBodyStepKernel::Name(), contextRef);
    kernel.getAs<IntegrateRigidBodyStepKernel>().initialize(contextRef, *this);
}

void RigidBodyIntegrator::cleanup() { kernel = rbhost::Kernel(); }

std::vector<std::string> RigidBodyIntegrator::getKernelNames() { return {IntegrateRigidBodyStepKernel::Name()}; }

// positions: forces, then geometry + dynamics; velocities: dynamics
// (RigidBodyIntegrator.cpp:63-74).  The platform has already rebuilt the body
// state from the new atom state with its device forces; the context's full
// force evaluation (host forces included) then redoes the first half-kick,
// and the host view refreshes its counts.
void RigidBodyIntegrator::stateChanged(rbhost::State::DataType changed) {
    if (changed == rbhost::State::Positions || changed == rbhost::State::Velocities) {
        if (changed == rbhost::State::Positions) context->calcForcesAndEnergy(true, false);
        auto& k = kernel.getAs<IntegrateRigidBodyStepKernel>();
        k.uploadBodySystem(bodySystem);
        bodySystem.update(*context, changed == rbhost::State::Positions, true);
    }
}

double RigidBodyIntegrator::computeKineticEnergy() {
    return kernel.getAs<IntegrateRigidBodyStepKernel>().computeKineticEnergy(*context, *this);
}

std::vector<double> RigidBodyIntegrator::getKineticEnergies() {
    if (context == nullptr) throw Exception("This Integrator is not bound to a context!");
    return kernel.getAs<IntegrateRigidBodyStepKernel>().getKineticEnergies(*this);
}

std::vector<double> RigidBodyIntegrator::getRefinedKineticEnergies() {
    if (context == nullptr) throw Exception("This Integrator is not bound to a context!");
    return kernel.getAs<IntegrateRigidBodyStepKernel>().getRefinedKineticEnergies(*this);
}

double RigidBodyIntegrator::getPotentialEnergyRefinement() {
    if (context == nullptr) throw Exception("This Integrator is not bound to a context!");
    return kernel.getAs<IntegrateRigidBodyStepKernel>().getPotentialEnergyRefinement(*this);
}

// RigidBodyIntegrator.cpp:96-101: one execute() per step, or all of them in
// one call of a kernel that can run them without the host in between
void RigidBodyIntegrator::step(int steps) {
    if (context == nullptr) throw Exception("This Integrator is not bound to a context!");
    if (steps <= 0) return;
    auto& k = kernel.getAs<IntegrateRigidBodyStepKernel>();
    auto* multi = dynamic_cast<MultiStepKernel*>(&k);
    if (multi != nullptr && multi->canExecuteSteps(*context)) {
        multi->executeSteps(*context, *this, steps);
        return;
    }
    for (int i = 0; i < steps; i++) k.execute(*context, *this);
}

}  // namespace RigidBodyPlugin
