// The HIP platform of the host API (see HipPlatform.h).
#include "HipPlatform.h"

#include <cstdlib>

namespace rbhost {

void rbhip_check(int status) {
    if (status != RBHIP_OK) throw Exception(rbhip_last_error());
}

namespace {
std::vector<double> flatten(const std::vector<Vec3>& v) {
    std::vector<double> out(3 * v.size());
    for (size_t i = 0; i < v.size(); i++)
        for (int c = 0; c < 3; c++) out[3 * i + c] = v[i][c];
    return out;
}
void unflatten(const std::vector<double>& in, std::vector<Vec3>& v) {
    v.resize(in.size() / 3);
    for (size_t i = 0; i < v.size(); i++) v[i] = Vec3(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
}
}  // namespace

HipPlatform::PlatformData::~PlatformData() {
    if (sim) (void)rbhip_destroy(sim);
}

rbhip_context* HipPlatform::PlatformData::require() const {
    if (!sim) throw Exception("The HIP platform holds no simulation state: the Context's integrator has not created it");
    return sim;
}

void HipPlatform::PlatformData::createSimulation(const System& system, const std::vector<int>& bodyIndices,
                                                 double stepSize, int rotationMode) {
    if (sim) {
        (void)rbhip_destroy(sim);
        sim = nullptr;
    }
    const int n = system.getNumParticles();
    std::vector<double> masses(n);
    for (int i = 0; i < n; i++) masses[i] = system.getParticleMass(i);
    const int nc = system.getNumConstraints();
    std::vector<int> pairs(2 * (size_t)nc);
    std::vector<double> dist(nc);
    for (int c = 0; c < nc; c++) system.getConstraintParameters(c, pairs[2 * c], pairs[2 * c + 1], dist[c]);
    rbhip_context* s = nullptr;
    rbhip_check(rbhip_create_constrained(n, masses.data(), bodyIndices.data(), nc, pairs.data(), dist.data(),
                                         stepSize, rotationMode, precision, device, &s));
    try {
        for (int f = 0; f < system.getNumForces(); f++) {
            const Force& force = system.getForce(f);
            if (auto* bonds = dynamic_cast<const HarmonicBondForce*>(&force)) {
                const int nb = bonds->getNumBonds();
                std::vector<int> pairs(2 * (size_t)nb);
                std::vector<double> len(nb), k(nb);
                for (int b = 0; b < nb; b++) bonds->getBondParameters(b, pairs[2 * b], pairs[2 * b + 1], len[b], k[b]);
                rbhip_check(rbhip_add_bonds(s, nb, pairs.data(), len.data(), k.data()));
            } else if (auto* tether = dynamic_cast<const TetherForce*>(&force)) {
                std::vector<double> k(n, 0.0), r0(3 * (size_t)n, 0.0);
                for (int t = 0; t < tether->getNumParticles(); t++) {
                    int atom;
                    double kt;
                    Vec3 r;
                    tether->getParticleParameters(t, atom, kt, r);
                    if (atom < 0 || atom >= n) throw Exception("TetherForce particle index out of range");
                    k[atom] = kt;
                    for (int c = 0; c < 3; c++) r0[3 * (size_t)atom + c] = r[c];
                }
                rbhip_check(rbhip_set_tether(s, k.data(), r0.data()));
            } else if (dynamic_cast<const HostForce*>(&force) == nullptr) {
                throw Exception("The HIP platform supports HarmonicBondForce, TetherForce and HostForce only");
            }
        }
        std::vector<int> types, siteAtoms, parents;
        std::vector<double> weights;
        for (int i = 0; i < n; i++) {
            if (!system.isVirtualSite(i)) continue;
            const VirtualSite& v = system.getVirtualSite(i);
            double w[12];
            v.weights(w);
            types.push_back(v.type());
            siteAtoms.push_back(i);
            for (int k = 0; k < 3; k++) parents.push_back(k < v.getNumParticles() ? v.getParticle(k) : 0);
            weights.insert(weights.end(), w, w + 12);
        }
        if (!types.empty())
            rbhip_check(rbhip_set_virtual_sites(s, (int)types.size(), types.data(), siteAtoms.data(), parents.data(),
                                                weights.data()));
    } catch (...) {
        (void)rbhip_destroy(s);
        throw;
    }
    sim = s;
}

bool HipPlatform::PlatformData::hasHostForces(const System& system) const {
    for (int f = 0; f < system.getNumForces(); f++)
        if (dynamic_cast<const HostForce*>(&system.getForce(f))) return true;
    return false;
}

HipPlatform::HipPlatform() = default;

HipPlatform::PlatformData& HipPlatform::data(ContextImpl& context) {
    auto* d = static_cast<PlatformData*>(context.getPlatformData());
    if (!d) throw Exception("Context has no HIP platform data");
    return *d;
}

void HipPlatform::contextCreated(ContextImpl& context, const std::map<std::string, std::string>& properties) const {
    auto* d = new PlatformData();
    try {
        for (const auto& kv : properties) {
            if (kv.first == "DeviceIndex") {
                char* end = nullptr;
                long v = std::strtol(kv.second.c_str(), &end, 10);
                if (kv.second.empty() || *end != '\0' || v < 0) throw Exception("Illegal value for DeviceIndex: " + kv.second);
                d->device = (int)v;
            } else if (kv.first == "Precision") {
                if (kv.second == "mixed") d->precision = RBHIP_PRECISION_MIXED;
                else if (kv.second == "double") d->precision = RBHIP_PRECISION_DOUBLE;
                else throw Exception("Illegal value for Precision: " + kv.second + " (mixed or double)");
            } else {
                throw Exception("Illegal property name for the HIP platform: " + kv.first);
            }
        }
    } catch (...) {
        delete d;
        throw;
    }
    context.setPlatformData(d);
}

void HipPlatform::contextDestroyed(ContextImpl& context) const {
    delete static_cast<PlatformData*>(context.getPlatformData());
    context.setPlatformData(nullptr);
}

void HipPlatform::setPositions(ContextImpl& context, const std::vector<Vec3>& positions) const {
    std::vector<double> xyz = flatten(positions);
    rbhip_check(rbhip_set_positions(data(context).require(), xyz.data()));
}

void HipPlatform::setVelocities(ContextImpl& context, const std::vector<Vec3>& velocities) const {
    std::vector<double> xyz = flatten(velocities);
    rbhip_check(rbhip_set_velocities(data(context).require(), xyz.data()));
}

void HipPlatform::getPositions(ContextImpl& context, std::vector<Vec3>& positions) const {
    std::vector<double> xyz(3 * (size_t)context.getSystem().getNumParticles());
    rbhip_check(rbhip_get_positions(data(context).require(), xyz.data()));
    unflatten(xyz, positions);
}

void HipPlatform::getVelocities(ContextImpl& context, std::vector<Vec3>& velocities) const {
    std::vector<double> xyz(3 * (size_t)context.getSystem().getNumParticles());
    rbhip_check(rbhip_get_velocities(data(context).require(), xyz.data()));
    unflatten(xyz, velocities);
}

void HipPlatform::getForces(ContextImpl& context, std::vector<Vec3>& forces) const {
    std::vector<double> xyz(3 * (size_t)context.getSystem().getNumParticles());
    rbhip_check(rbhip_get_forces(data(context).require(), xyz.data()));
    unflatten(xyz, forces);
}

// the device force field, then each HostForce at the current positions,
// added to the device forces and handed back through rbhip_set_forces
double HipPlatform::calcForcesAndEnergy(ContextImpl& context, bool includeEnergy) const {
    PlatformData& d = data(context);
    rbhip_context* sim = d.require();
    const System& system = context.getSystem();
    double e = 0.0;
    const bool host = d.hasHostForces(system);
    rbhip_check(rbhip_compute_forces(sim, (includeEnergy || host) ? &e : nullptr));
    if (!host) return e;
    const size_t n = (size_t)system.getNumParticles();
    std::vector<double> xyz(3 * n), f(3 * n);
    rbhip_check(rbhip_get_positions(sim, xyz.data()));
    rbhip_check(rbhip_get_forces(sim, f.data()));
    std::vector<Vec3> positions, forces;
    unflatten(xyz, positions);
    unflatten(f, forces);
    for (int k = 0; k < system.getNumForces(); k++)
        if (auto* hf = dynamic_cast<const HostForce*>(&system.getForce(k))) {
            std::vector<Vec3> fk(n);
            e += hf->evaluate(positions, fk);
            for (size_t i = 0; i < n; i++) forces[i] = forces[i] + fk[i];
        }
    f = flatten(forces);
    rbhip_check(rbhip_set_forces(sim, f.data()));
    return e;
}

double HipPlatform::getTime(ContextImpl& context) const {
    double t = 0.0;
    rbhip_check(rbhip_get_time(data(context).require(), &t, nullptr));
    return t;
}

long long HipPlatform::getStepCount(ContextImpl& context) const {
    long long n = 0;
    rbhip_check(rbhip_get_time(data(context).require(), nullptr, &n));
    return n;
}

}  // namespace rbhost
