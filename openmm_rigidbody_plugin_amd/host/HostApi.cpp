// Host-side simulation API (see HostApi.h).
#include "HostApi.h"

#include <algorithm>

#include "rbhip.h"

namespace rbhost {

namespace {
std::vector<Platform*>& platforms() {
    static std::vector<Platform*> list;
    return list;
}
void check_index(int index, int size, const char* what) {
    if (index < 0 || index >= size) throw Exception(std::string(what) + " index out of range");
}
}  // namespace

// ---------------------------------------------------------------------------
int HarmonicBondForce::addBond(int particle1, int particle2, double length, double k) {
    p1.push_back(particle1);
    p2.push_back(particle2);
    len.push_back(length);
    kb.push_back(k);
    return (int)p1.size() - 1;
}

void HarmonicBondForce::getBondParameters(int index, int& particle1, int& particle2, double& length,
                                          double& k) const {
    check_index(index, (int)p1.size(), "Bond");
    particle1 = p1[index];
    particle2 = p2[index];
    length = len[index];
    k = kb[index];
}

int TetherForce::addParticle(int particle, double k, const Vec3& r0) {
    atom.push_back(particle);
    kk.push_back(k);
    anchor.push_back(r0);
    return (int)atom.size() - 1;
}

void TetherForce::getParticleParameters(int index, int& particle, double& k, Vec3& r0) const {
    check_index(index, (int)atom.size(), "Tether");
    particle = atom[index];
    k = kk[index];
    r0 = anchor[index];
}

// ---------------------------------------------------------------------------
int VirtualSite::getParticle(int index) const {
    check_index(index, (int)particles.size(), "Virtual site particle");
    return particles[index];
}

TwoParticleAverageSite::TwoParticleAverageSite(int particle1, int particle2, double weight1, double weight2)
    : w1(weight1), w2(weight2) {
    particles = {particle1, particle2};
}
double TwoParticleAverageSite::getWeight(int particle) const {
    check_index(particle, 2, "Virtual site particle");
    return particle == 0 ? w1 : w2;
}
int TwoParticleAverageSite::type() const { return RBHIP_SITE_TWO_AVERAGE; }
void TwoParticleAverageSite::weights(double w[12]) const {
    for (int k = 0; k < 12; k++) w[k] = 0.0;
    w[0] = w1;
    w[1] = w2;
}

ThreeParticleAverageSite::ThreeParticleAverageSite(int particle1, int particle2, int particle3, double weight1,
                                                   double weight2, double weight3)
    : w{weight1, weight2, weight3} {
    particles = {particle1, particle2, particle3};
}
double ThreeParticleAverageSite::getWeight(int particle) const {
    check_index(particle, 3, "Virtual site particle");
    return w[particle];
}
int ThreeParticleAverageSite::type() const { return RBHIP_SITE_THREE_AVERAGE; }
void ThreeParticleAverageSite::weights(double out[12]) const {
    for (int k = 0; k < 12; k++) out[k] = k < 3 ? w[k] : 0.0;
}

OutOfPlaneSite::OutOfPlaneSite(int particle1, int particle2, int particle3, double weight12, double weight13,
                               double weightCross)
    : w12(weight12), w13(weight13), wc(weightCross) {
    particles = {particle1, particle2, particle3};
}
int OutOfPlaneSite::type() const { return RBHIP_SITE_OUT_OF_PLANE; }
void OutOfPlaneSite::weights(double w[12]) const {
    for (int k = 0; k < 12; k++) w[k] = 0.0;
    w[0] = w12;
    w[1] = w13;
    w[2] = wc;
}

LocalCoordinatesSite::LocalCoordinatesSite(int particle1, int particle2, int particle3, const Vec3& originWeights,
                                           const Vec3& xWeights, const Vec3& yWeights, const Vec3& localPosition)
    : ow(originWeights), xw(xWeights), yw(yWeights), lp(localPosition) {
    particles = {particle1, particle2, particle3};
}
int LocalCoordinatesSite::type() const { return RBHIP_SITE_LOCAL_COORDINATES; }
void LocalCoordinatesSite::weights(double w[12]) const {
    for (int k = 0; k < 3; k++) {
        w[k] = ow[k];
        w[3 + k] = xw[k];
        w[6 + k] = yw[k];
        w[9 + k] = lp[k];
    }
}

int System::addParticle(double mass) {
    masses.push_back(mass);
    sites.emplace_back();
    return (int)masses.size() - 1;
}

void System::setVirtualSite(int index, VirtualSite* site) {
    check_index(index, (int)masses.size(), "Particle");
    sites[index].reset(site);
}

bool System::isVirtualSite(int index) const {
    check_index(index, (int)masses.size(), "Particle");
    return sites[index] != nullptr;
}

const VirtualSite& System::getVirtualSite(int index) const {
    if (!isVirtualSite(index)) throw Exception("This particle is not a virtual site");
    return *sites[index];
}

double System::getParticleMass(int index) const {
    check_index(index, (int)masses.size(), "Particle");
    return masses[index];
}

void System::setParticleMass(int index, double mass) {
    check_index(index, (int)masses.size(), "Particle");
    masses[index] = mass;
}

int System::addConstraint(int particle1, int particle2, double distance) {
    c1.push_back(particle1);
    c2.push_back(particle2);
    cd.push_back(distance);
    return (int)c1.size() - 1;
}

void System::getConstraintParameters(int index, int& particle1, int& particle2, double& distance) const {
    check_index(index, (int)c1.size(), "Constraint");
    particle1 = c1[index];
    particle2 = c2[index];
    distance = cd[index];
}

int System::addForce(Force* force) {
    forces.emplace_back(force);
    return (int)forces.size() - 1;
}

Force& System::getForce(int index) {
    check_index(index, (int)forces.size(), "Force");
    return *forces[index];
}

const Force& System::getForce(int index) const {
    check_index(index, (int)forces.size(), "Force");
    return *forces[index];
}

// ---------------------------------------------------------------------------
const std::vector<Vec3>& State::getPositions() const {
    if (!(types & Positions)) throw Exception("Invoked getPositions() on a State which does not contain positions.");
    return positions;
}

const std::vector<Vec3>& State::getVelocities() const {
    if (!(types & Velocities))
        throw Exception("Invoked getVelocities() on a State which does not contain velocities.");
    return velocities;
}

const std::vector<Vec3>& State::getForces() const {
    if (!(types & Forces)) throw Exception("Invoked getForces() on a State which does not contain forces.");
    return forces;
}

double State::getKineticEnergy() const {
    if (!(types & Energy)) throw Exception("Invoked getKineticEnergy() on a State which does not contain energies.");
    return kinetic;
}

double State::getPotentialEnergy() const {
    if (!(types & Energy)) throw Exception("Invoked getPotentialEnergy() on a State which does not contain energies.");
    return potential;
}

// ---------------------------------------------------------------------------
const std::string& Kernel::getName() const {
    if (!impl) throw Exception("Kernel has not been initialized");
    return impl->getName();
}

void Platform::registerPlatform(Platform* platform) { platforms().push_back(platform); }

int Platform::getNumPlatforms() { return (int)platforms().size(); }

Platform& Platform::getPlatform(int index) {
    check_index(index, (int)platforms().size(), "Platform");
    return *platforms()[index];
}

Platform& Platform::getPlatformByName(const std::string& name) {
    for (Platform* p : platforms())
        if (p->getName() == name) return *p;
    throw Exception("There is no registered Platform called \"" + name + "\"");
}

void Platform::registerKernelFactory(const std::string& name, KernelFactory* factory) {
    factories[name] = std::shared_ptr<KernelFactory>(factory);
}

bool Platform::supportsKernels(const std::vector<std::string>& names) const {
    for (const std::string& n : names)
        if (factories.find(n) == factories.end()) return false;
    return true;
}

Kernel Platform::createKernel(const std::string& name, ContextImpl& context) const {
    auto it = factories.find(name);
    if (it == factories.end()) throw Exception("Called createKernel() on a Platform which does not support the requested kernel");
    return Kernel(it->second->createKernelImpl(name, *this, context));
}

// ---------------------------------------------------------------------------
void Integrator::setStepSize(double size) {
    if (!(size > 0.0)) throw Exception("Step size must be positive");
    stepSize = size;
}

// ---------------------------------------------------------------------------
ContextImpl::ContextImpl(Context& owner, System& system, Integrator& integrator, Platform& platform,
                         const std::map<std::string, std::string>& properties)
    : owner(owner), system(system), integrator(integrator), platform(platform) {
    if (system.getNumParticles() == 0) throw Exception("Cannot create a Context for a System with no particles");
    if (integrator.owner != nullptr) throw Exception("This Integrator is already bound to a context");
    if (!platform.supportsKernels(integrator.getKernelNames()))
        throw Exception("Specified a Platform for a Context which does not support all required kernels");
    platform.contextCreated(*this, properties);
    try {
        integrator.initialize(*this);
    } catch (...) {
        integrator.cleanup();
        integrator.context = nullptr;
        integrator.owner = nullptr;
        platform.contextDestroyed(*this);
        throw;
    }
}

ContextImpl::~ContextImpl() {
    integrator.cleanup();
    integrator.context = nullptr;
    integrator.owner = nullptr;
    platform.contextDestroyed(*this);
}

void ContextImpl::setPositions(const std::vector<Vec3>& positions) {
    if ((int)positions.size() != system.getNumParticles())
        throw Exception("Called setPositions() on a Context with the wrong number of positions");
    platform.setPositions(*this, positions);
    integrator.stateChanged(State::Positions);
}

void ContextImpl::setVelocities(const std::vector<Vec3>& velocities) {
    if ((int)velocities.size() != system.getNumParticles())
        throw Exception("Called setVelocities() on a Context with the wrong number of velocities");
    platform.setVelocities(*this, velocities);
    integrator.stateChanged(State::Velocities);
}

void ContextImpl::getPositions(std::vector<Vec3>& positions) { platform.getPositions(*this, positions); }
void ContextImpl::getVelocities(std::vector<Vec3>& velocities) { platform.getVelocities(*this, velocities); }
void ContextImpl::getForces(std::vector<Vec3>& forces) { platform.getForces(*this, forces); }

double ContextImpl::calcForcesAndEnergy(bool, bool includeEnergy) {
    return platform.calcForcesAndEnergy(*this, includeEnergy);
}

double ContextImpl::calcKineticEnergy() { return integrator.computeKineticEnergy(); }
double ContextImpl::getTime() { return platform.getTime(*this); }
long long ContextImpl::getStepCount() { return platform.getStepCount(*this); }

Context::Context(System& system, Integrator& integrator, Platform& platform,
                 const std::map<std::string, std::string>& properties)
    : impl(new ContextImpl(*this, system, integrator, platform, properties)) {}

Context::~Context() = default;

void Context::setPositions(const std::vector<Vec3>& positions) { impl->setPositions(positions); }
void Context::setVelocities(const std::vector<Vec3>& velocities) { impl->setVelocities(velocities); }

State Context::getState(int types) const {
    State s;
    s.types = types;
    s.time = impl->getTime();
    s.stepCount = impl->getStepCount();
    if (types & State::Positions) impl->getPositions(s.positions);
    if (types & State::Velocities) impl->getVelocities(s.velocities);
    if (types & State::Energy) {
        s.potential = impl->calcForcesAndEnergy(false, true);
        s.kinetic = impl->calcKineticEnergy();
    }
    if (types & State::Forces) impl->getForces(s.forces);
    return s;
}

}  // namespace rbhost
