// The "HIP" platform of the host API: each Context's atom state (positions,
// velocities, forces), its force field and the rigid-body integrator state
// live in one rbhip context on one gfx950 device.
//
// Platform properties (Context constructor):
//   "DeviceIndex"  HIP device ordinal, default "0"
//   "Precision"    "mixed" (fp32 state, FP64 accumulation; default) or "double"
#pragma once

#include <string>

#include "HostApi.h"
#include "rbhip.h"

namespace rbhost {

class HipPlatform : public Platform {
public:
    class PlatformData {
    public:
        int device = 0;
        int precision = RBHIP_PRECISION_MIXED;
        rbhip_context* sim = nullptr;   // created by the integrator's kernel
        ~PlatformData();
        // creates the device context for `system` and installs its device
        // forces (HarmonicBondForce, TetherForce); HostForce is evaluated on
        // the host at each force evaluation; anything else is rejected
        void createSimulation(const System& system, const std::vector<int>& bodyIndices, double stepSize,
                              int rotationMode);
        rbhip_context* require() const;
        bool hasHostForces(const System& system) const;
    };

    HipPlatform();
    const std::string& getName() const override { return name; }
    void contextCreated(ContextImpl& context, const std::map<std::string, std::string>& properties) const override;
    void contextDestroyed(ContextImpl& context) const override;
    void setPositions(ContextImpl& context, const std::vector<Vec3>& positions) const override;
    void setVelocities(ContextImpl& context, const std::vector<Vec3>& velocities) const override;
    void getPositions(ContextImpl& context, std::vector<Vec3>& positions) const override;
    void getVelocities(ContextImpl& context, std::vector<Vec3>& velocities) const override;
    void getForces(ContextImpl& context, std::vector<Vec3>& forces) const override;
    double calcForcesAndEnergy(ContextImpl& context, bool includeEnergy) const override;
    double getTime(ContextImpl& context) const override;
    long long getStepCount(ContextImpl& context) const override;

    static PlatformData& data(ContextImpl& context);

private:
    std::string name = "HIP";
};

// throws rbhost::Exception with the library's message unless status is RBHIP_OK
void rbhip_check(int status);

}  // namespace rbhost
