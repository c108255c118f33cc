// Minimal host-side simulation API for the HIP rigid-body plugin.
//
// It mirrors the subset of OpenMM's public API the RigidBodyIntegrator path
// touches (System, Context, State, Integrator, Platform, KernelImpl,
// KernelFactory, HarmonicBondForce), with the same names, argument meaning
// and exception behaviour, so the plugin classes and their tests read like
// the reference's (tests/TestRigidBodyIntegrator.h).  Everything here is
// host C++; device work goes through the rbhip C ABI (include/rbhip.h).
//
// Force fields: HarmonicBondForce and TetherForce (a per-particle harmonic
// restraint, E = k/2 |r - r0|^2, the benchmark's stand-in force field) are
// evaluated on the device inside the rbhip context; HostForce is any other
// force field, evaluated on the host by a callback and handed to the
// integrator through the same force slot.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rbhost {

class Exception : public std::runtime_error {
public:
    explicit Exception(const std::string& message) : std::runtime_error(message) {}
};

struct Vec3 {
    double v[3];
    Vec3() : v{0.0, 0.0, 0.0} {}
    Vec3(double x, double y, double z) : v{x, y, z} {}
    double operator[](int i) const { return v[i]; }
    double& operator[](int i) { return v[i]; }
    Vec3 operator+(const Vec3& o) const { return Vec3(v[0] + o[0], v[1] + o[1], v[2] + o[2]); }
    Vec3 operator-(const Vec3& o) const { return Vec3(v[0] - o[0], v[1] - o[1], v[2] - o[2]); }
    Vec3 operator*(double s) const { return Vec3(v[0] * s, v[1] * s, v[2] * s); }
    double dot(const Vec3& o) const { return v[0] * o[0] + v[1] * o[1] + v[2] * o[2]; }
    Vec3 cross(const Vec3& o) const {
        return Vec3(v[1] * o[2] - v[2] * o[1], v[2] * o[0] - v[0] * o[2], v[0] * o[1] - v[1] * o[0]);
    }
};

// ---------------------------------------------------------------------------
// forces
// ---------------------------------------------------------------------------
class Force {
public:
    virtual ~Force() = default;
};

class HarmonicBondForce : public Force {
public:
    int addBond(int particle1, int particle2, double length, double k);
    int getNumBonds() const { return (int)p1.size(); }
    void getBondParameters(int index, int& particle1, int& particle2, double& length, double& k) const;

private:
    std::vector<int> p1, p2;
    std::vector<double> len, kb;
};

// A force field evaluated on the host: the callback receives the positions
// of all particles (FP64) and adds its forces to `forces` (same length,
// zero-initialised by the caller) and returns its potential energy.  It
// stands for any force field the platform does not evaluate itself.
class HostForce : public Force {
public:
    using Callback = std::function<double(const std::vector<Vec3>& positions, std::vector<Vec3>& forces)>;
    explicit HostForce(Callback callback) : callback(std::move(callback)) {}
    double evaluate(const std::vector<Vec3>& positions, std::vector<Vec3>& forces) const {
        return callback(positions, forces);
    }

private:
    Callback callback;
};

class TetherForce : public Force {
public:
    // restrains `particle` to r0 with spring constant k (kJ/mol/nm^2)
    int addParticle(int particle, double k, const Vec3& r0);
    int getNumParticles() const { return (int)atom.size(); }
    void getParticleParameters(int index, int& particle, double& k, Vec3& r0) const;

private:
    std::vector<int> atom;
    std::vector<double> kk;
    std::vector<Vec3> anchor;
};

// ---------------------------------------------------------------------------
// System
// ---------------------------------------------------------------------------
// OpenMM's virtual sites of the four kinds this platform computes; a site
// particle must be massless.  type() is the RBHIP_SITE_* value.
class VirtualSite {
public:
    virtual ~VirtualSite() = default;
    int getNumParticles() const { return (int)particles.size(); }
    int getParticle(int index) const;
    virtual int type() const = 0;
    // 12 values, as rbhip_set_virtual_sites takes them: {w1, w2}, {w1, w2, w3},
    // {w12, w13, wCross}, or origin, x, y weights and the local position
    virtual void weights(double w[12]) const = 0;

protected:
    std::vector<int> particles;
};

class TwoParticleAverageSite : public VirtualSite {
public:
    TwoParticleAverageSite(int particle1, int particle2, double weight1, double weight2);
    double getWeight(int particle) const;
    int type() const override;
    void weights(double w[12]) const override;

private:
    double w1, w2;
};

class ThreeParticleAverageSite : public VirtualSite {
public:
    ThreeParticleAverageSite(int particle1, int particle2, int particle3, double weight1, double weight2,
                             double weight3);
    double getWeight(int particle) const;
    int type() const override;
    void weights(double w[12]) const override;

private:
    double w[3];
};

class OutOfPlaneSite : public VirtualSite {
public:
    OutOfPlaneSite(int particle1, int particle2, int particle3, double weight12, double weight13, double weightCross);
    double getWeight12() const { return w12; }
    double getWeight13() const { return w13; }
    double getWeightCross() const { return wc; }
    int type() const override;
    void weights(double w[12]) const override;

private:
    double w12, w13, wc;
};

class LocalCoordinatesSite : public VirtualSite {
public:
    LocalCoordinatesSite(int particle1, int particle2, int particle3, const Vec3& originWeights, const Vec3& xWeights,
                         const Vec3& yWeights, const Vec3& localPosition);
    Vec3 getOriginWeights() const { return ow; }
    Vec3 getXWeights() const { return xw; }
    Vec3 getYWeights() const { return yw; }
    Vec3 getLocalPosition() const { return lp; }
    int type() const override;
    void weights(double w[12]) const override;

private:
    Vec3 ow, xw, yw, lp;
};

class System {
public:
    System() = default;
    System(const System&) = delete;
    System& operator=(const System&) = delete;
    int addParticle(double mass);
    int getNumParticles() const { return (int)masses.size(); }
    double getParticleMass(int index) const;
    void setParticleMass(int index, double mass);
    // takes ownership; the particle should be massless
    void setVirtualSite(int index, VirtualSite* site);
    bool isVirtualSite(int index) const;
    const VirtualSite& getVirtualSite(int index) const;
    int addConstraint(int particle1, int particle2, double distance);
    int getNumConstraints() const { return (int)c1.size(); }
    void getConstraintParameters(int index, int& particle1, int& particle2, double& distance) const;
    // takes ownership
    int addForce(Force* force);
    int getNumForces() const { return (int)forces.size(); }
    Force& getForce(int index);
    const Force& getForce(int index) const;

private:
    std::vector<double> masses;
    std::vector<std::unique_ptr<VirtualSite>> sites;   // per particle, or null
    std::vector<int> c1, c2;
    std::vector<double> cd;
    std::vector<std::unique_ptr<Force>> forces;
};

// ---------------------------------------------------------------------------
// State
// ---------------------------------------------------------------------------
class State {
public:
    enum DataType { Positions = 1, Velocities = 2, Forces = 4, Energy = 8 };
    double getTime() const { return time; }
    long long getStepCount() const { return stepCount; }
    const std::vector<Vec3>& getPositions() const;
    const std::vector<Vec3>& getVelocities() const;
    const std::vector<Vec3>& getForces() const;
    double getKineticEnergy() const;
    double getPotentialEnergy() const;

private:
    friend class Context;
    int types = 0;
    double time = 0.0;
    long long stepCount = 0;
    std::vector<Vec3> positions, velocities, forces;
    double kinetic = 0.0, potential = 0.0;
};

class Context;
class ContextImpl;
class Platform;

// ---------------------------------------------------------------------------
// kernels and platforms
// ---------------------------------------------------------------------------
class KernelImpl {
public:
    KernelImpl(const std::string& name, const Platform& platform) : name(name), platform(&platform) {}
    virtual ~KernelImpl() = default;
    const std::string& getName() const { return name; }
    const Platform& getPlatform() const { return *platform; }

private:
    std::string name;
    const Platform* platform;
};

class Kernel {
public:
    Kernel() = default;
    explicit Kernel(KernelImpl* impl) : impl(impl) {}
    const std::string& getName() const;
    template <class T> T& getAs() {
        T* t = dynamic_cast<T*>(impl.get());
        if (!t) throw Exception("Kernel is not of the requested type");
        return *t;
    }
    bool isValid() const { return impl != nullptr; }

private:
    std::shared_ptr<KernelImpl> impl;
};

class KernelFactory {
public:
    virtual ~KernelFactory() = default;
    virtual KernelImpl* createKernelImpl(std::string name, const Platform& platform, ContextImpl& context) const = 0;
};

// Platform: owns its per-context data and the atom state of a context.
class Platform {
public:
    virtual ~Platform() = default;
    virtual const std::string& getName() const = 0;
    // registry (a platform is registered once and lives until exit)
    static void registerPlatform(Platform* platform);
    static int getNumPlatforms();
    static Platform& getPlatform(int index);
    static Platform& getPlatformByName(const std::string& name);
    // kernel factories; the platform takes ownership of each factory
    void registerKernelFactory(const std::string& name, KernelFactory* factory);
    bool supportsKernels(const std::vector<std::string>& names) const;
    Kernel createKernel(const std::string& name, ContextImpl& context) const;
    // context lifetime and the context's atom state
    virtual void contextCreated(ContextImpl& context, const std::map<std::string, std::string>& properties) const = 0;
    virtual void contextDestroyed(ContextImpl& context) const = 0;
    virtual void setPositions(ContextImpl& context, const std::vector<Vec3>& positions) const = 0;
    virtual void setVelocities(ContextImpl& context, const std::vector<Vec3>& velocities) const = 0;
    virtual void getPositions(ContextImpl& context, std::vector<Vec3>& positions) const = 0;
    virtual void getVelocities(ContextImpl& context, std::vector<Vec3>& velocities) const = 0;
    virtual void getForces(ContextImpl& context, std::vector<Vec3>& forces) const = 0;
    // forces (and the potential energy if requested) at the current positions
    virtual double calcForcesAndEnergy(ContextImpl& context, bool includeEnergy) const = 0;
    virtual double getTime(ContextImpl& context) const = 0;
    virtual long long getStepCount(ContextImpl& context) const = 0;

private:
    std::map<std::string, std::shared_ptr<KernelFactory>> factories;
};

// ---------------------------------------------------------------------------
// Integrator
// ---------------------------------------------------------------------------
class Integrator {
public:
    virtual ~Integrator() = default;
    double getStepSize() const { return stepSize; }
    virtual void setStepSize(double size);
    double getConstraintTolerance() const { return constraintTol; }
    void setConstraintTolerance(double tol) { constraintTol = tol; }
    virtual void step(int steps) = 0;

protected:
    friend class ContextImpl;
    virtual void initialize(ContextImpl& context) = 0;
    virtual void cleanup() {}
    virtual std::vector<std::string> getKernelNames() = 0;
    virtual void stateChanged(State::DataType) {}
    virtual double computeKineticEnergy() = 0;
    ContextImpl* context = nullptr;
    Context* owner = nullptr;

private:
    double stepSize = 0.0;
    double constraintTol = 1e-5;
};

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
class ContextImpl {
public:
    ContextImpl(Context& owner, System& system, Integrator& integrator, Platform& platform,
                const std::map<std::string, std::string>& properties);
    ~ContextImpl();
    Context& getOwner() { return owner; }
    const System& getSystem() const { return system; }
    Integrator& getIntegrator() { return integrator; }
    Platform& getPlatform() { return platform; }
    void* getPlatformData() { return platformData; }
    void setPlatformData(void* data) { platformData = data; }
    void setPositions(const std::vector<Vec3>& positions);
    void setVelocities(const std::vector<Vec3>& velocities);
    void getPositions(std::vector<Vec3>& positions);
    void getVelocities(std::vector<Vec3>& velocities);
    void getForces(std::vector<Vec3>& forces);
    double calcForcesAndEnergy(bool includeForces, bool includeEnergy);
    double calcKineticEnergy();
    double getTime();
    long long getStepCount();
    // no-op here: positions are written straight to the platform state
    void updateContextState() {}

private:
    Context& owner;
    System& system;
    Integrator& integrator;
    Platform& platform;
    void* platformData = nullptr;
};

class Context {
public:
    Context(System& system, Integrator& integrator, Platform& platform,
            const std::map<std::string, std::string>& properties = {});
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    State getState(int types) const;
    void setPositions(const std::vector<Vec3>& positions);
    void setVelocities(const std::vector<Vec3>& velocities);
    const System& getSystem() const { return impl->getSystem(); }
    Integrator& getIntegrator() { return impl->getIntegrator(); }
    Platform& getPlatform() { return impl->getPlatform(); }

private:
    std::unique_ptr<ContextImpl> impl;
};

}  // namespace rbhost
