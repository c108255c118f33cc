// Plugin registration and the handle-based C entry points of the host
// plugin library (include/rigidbody_hip_plugin.h).
#include <memory>
#include <string>

#include "HipRigidBodyKernels.h"
#include "rigidbody_hip_plugin.h"

using namespace RigidBodyPlugin;
using rbhost::Exception;
using rbhost::Vec3;

extern "C" void registerPlatforms(void) {
    try {
        rbhost::Platform::getPlatformByName("HIP");
    } catch (const Exception&) {
        rbhost::Platform::registerPlatform(new rbhost::HipPlatform());
    }
}

extern "C" void registerKernelFactories(void) {
    try {
        rbhost::Platform& platform = rbhost::Platform::getPlatformByName("HIP");
        if (!platform.supportsKernels({IntegrateRigidBodyStepKernel::Name()}))
            platform.registerKernelFactory(IntegrateRigidBodyStepKernel::Name(), new HipRigidBodyKernelFactory());
    } catch (const Exception&) {
        // no HIP platform registered: nothing to attach to
    }
}

extern "C" void registerRigidBodyHipKernelFactories(void) {
    registerPlatforms();
    registerKernelFactories();
}

// ---------------------------------------------------------------------------
struct rbplugin_sim {
    rbhost::System system;
    std::unique_ptr<RigidBodyIntegrator> integrator;
    std::unique_ptr<rbhost::Context> context;
    ~rbplugin_sim() {
        context.reset();
        integrator.reset();
    }
};

namespace {
thread_local std::string g_err;

template <class F> int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void need(bool ok, const char* msg) {
    if (!ok) throw Exception(msg);
}

std::vector<Vec3> vecs(const double* xyz, int n) {
    std::vector<Vec3> v(n);
    for (int i = 0; i < n; i++) v[i] = Vec3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return v;
}

void flat(const std::vector<Vec3>& v, double* out) {
    for (size_t i = 0; i < v.size(); i++)
        for (int c = 0; c < 3; c++) out[3 * i + c] = v[i][c];
}
}  // namespace

extern "C" const char* rbplugin_last_error(void) { return g_err.c_str(); }

extern "C" int rbplugin_create(int numAtoms, const double* masses, const int* bodyIndices, double stepSize,
                               int rotationMode, int precision, int device, int numBonds, const int* bondPairs,
                               const double* bondLength, const double* bondK, const double* tetherK,
                               const double* tetherR0, rbplugin_sim** out) {
    return rbplugin_create_ex(numAtoms, masses, bodyIndices, stepSize, rotationMode, 0, precision, device, numBonds,
                              bondPairs, bondLength, bondK, tetherK, tetherR0, out);
}

extern "C" int rbplugin_create_ex(int numAtoms, const double* masses, const int* bodyIndices, double stepSize,
                                  int rotationMode, int computeRefinedEnergies, int precision, int device,
                                  int numBonds, const int* bondPairs, const double* bondLength, const double* bondK,
                                  const double* tetherK, const double* tetherR0, rbplugin_sim** out) {
    return rbplugin_create_with_force(numAtoms, masses, bodyIndices, stepSize, rotationMode, computeRefinedEnergies,
                                      precision, device, numBonds, bondPairs, bondLength, bondK, tetherK, tetherR0,
                                      nullptr, nullptr, out);
}

extern "C" int rbplugin_create_with_force(int numAtoms, const double* masses, const int* bodyIndices,
                                          double stepSize, int rotationMode, int computeRefinedEnergies,
                                          int precision, int device, int numBonds, const int* bondPairs,
                                          const double* bondLength, const double* bondK, const double* tetherK,
                                          const double* tetherR0, rbplugin_force_fn hostForce, void* user,
                                          rbplugin_sim** out) {
    return guard([&] {
        need(out != nullptr, "null output handle");
        *out = nullptr;
        need(numAtoms > 0 && masses && bodyIndices, "empty system");
        need(numBonds >= 0 && (numBonds == 0 || (bondPairs && bondLength && bondK)), "bad bond arguments");
        need(precision == 0 || precision == 1, "precision must be 0 (mixed) or 1 (double)");
        need(device >= 0, "negative device index");
        registerRigidBodyHipKernelFactories();
        std::unique_ptr<rbplugin_sim> s(new rbplugin_sim());
        for (int i = 0; i < numAtoms; i++) s->system.addParticle(masses[i]);
        if (numBonds > 0) {
            auto* bonds = new rbhost::HarmonicBondForce();
            for (int b = 0; b < numBonds; b++) bonds->addBond(bondPairs[2 * b], bondPairs[2 * b + 1], bondLength[b], bondK[b]);
            s->system.addForce(bonds);
        }
        if (tetherK) {
            need(tetherR0 != nullptr, "tether anchors missing");
            auto* tether = new rbhost::TetherForce();
            for (int i = 0; i < numAtoms; i++)
                if (tetherK[i] != 0.0)
                    tether->addParticle(i, tetherK[i], Vec3(tetherR0[3 * i], tetherR0[3 * i + 1], tetherR0[3 * i + 2]));
            s->system.addForce(tether);
        }
        if (hostForce) {
            s->system.addForce(new rbhost::HostForce(
                [hostForce, user](const std::vector<Vec3>& x, std::vector<Vec3>& f) {
                    const int n = (int)x.size();
                    std::vector<double> xyz(3 * (size_t)n), fxyz(3 * (size_t)n, 0.0);
                    flat(x, xyz.data());
                    double e = hostForce(user, n, xyz.data(), fxyz.data());
                    for (int i = 0; i < n; i++) f[i] = f[i] + Vec3(fxyz[3 * i], fxyz[3 * i + 1], fxyz[3 * i + 2]);
                    return e;
                }));
        }
        s->integrator.reset(new RigidBodyIntegrator(stepSize, std::vector<int>(bodyIndices, bodyIndices + numAtoms)));
        s->integrator->setRotationMode(rotationMode);
        s->integrator->setComputeRefinedEnergies(computeRefinedEnergies != 0);
        std::map<std::string, std::string> props = {{"DeviceIndex", std::to_string(device)},
                                                    {"Precision", precision == 0 ? "mixed" : "double"}};
        s->context.reset(new rbhost::Context(s->system, *s->integrator, rbhost::Platform::getPlatformByName("HIP"), props));
        *out = s.release();
    });
}

extern "C" int rbplugin_destroy(rbplugin_sim* sim) {
    return guard([&] { delete sim; });
}

extern "C" int rbplugin_set_positions(rbplugin_sim* sim, const double* xyz) {
    return guard([&] {
        need(sim && xyz, "null argument");
        sim->context->setPositions(vecs(xyz, sim->system.getNumParticles()));
    });
}

extern "C" int rbplugin_set_velocities(rbplugin_sim* sim, const double* xyz) {
    return guard([&] {
        need(sim && xyz, "null argument");
        sim->context->setVelocities(vecs(xyz, sim->system.getNumParticles()));
    });
}

extern "C" int rbplugin_step(rbplugin_sim* sim, int steps) {
    return guard([&] {
        need(sim, "null argument");
        sim->integrator->step(steps);
    });
}

extern "C" int rbplugin_set_step_size(rbplugin_sim* sim, double stepSize) {
    return guard([&] {
        need(sim, "null argument");
        sim->integrator->setStepSize(stepSize);
    });
}

extern "C" int rbplugin_get_state(rbplugin_sim* sim, double* positions, double* velocities, double* forces,
                                  double* energies3) {
    return guard([&] {
        need(sim, "null argument");
        int types = (positions ? rbhost::State::Positions : 0) | (velocities ? rbhost::State::Velocities : 0) |
                    (forces ? rbhost::State::Forces : 0) | (energies3 ? rbhost::State::Energy : 0);
        rbhost::State st = sim->context->getState(types);
        if (positions) flat(st.getPositions(), positions);
        if (velocities) flat(st.getVelocities(), velocities);
        if (forces) flat(st.getForces(), forces);
        if (energies3) {
            energies3[0] = st.getKineticEnergy();
            energies3[1] = st.getPotentialEnergy();
            energies3[2] = st.getTime();
        }
    });
}

extern "C" int rbplugin_energies(rbplugin_sim* sim, double* out5) {
    return guard([&] {
        need(sim && out5, "null argument");
        std::vector<double> ke = sim->integrator->getKineticEnergies();
        std::vector<double> rk = sim->integrator->getRefinedKineticEnergies();
        out5[0] = ke[0];
        out5[1] = ke[1];
        out5[2] = rk[0];
        out5[3] = rk[1];
        out5[4] = sim->integrator->getPotentialEnergyRefinement();
    });
}

extern "C" int rbplugin_counts(rbplugin_sim* sim, int* out5) {
    return guard([&] {
        need(sim && out5, "null argument");
        const RigidBodySystem& b = sim->integrator->getRigidBodySystem();
        out5[0] = b.getNumFree();
        out5[1] = b.getNumBodies();
        out5[2] = b.getNumActualAtoms();
        out5[3] = b.getNumBodyAtoms();
        out5[4] = b.getNumDOF();
    });
}

extern "C" int rbplugin_atom_index(rbplugin_sim* sim, int* out) {
    return guard([&] {
        need(sim && out, "null argument");
        const RigidBodySystem& b = sim->integrator->getRigidBodySystem();
        for (int i = 0; i < b.getNumActualAtoms(); i++) out[i] = b.getAtomIndex(i);
    });
}

extern "C" int rbplugin_rigid_body(rbplugin_sim* sim, int index, double* out33) {
    return guard([&] {
        need(sim && out33, "null argument");
        RigidBody r = sim->integrator->getRigidBodySystem().getRigidBody(index);
        double* o = out33;
        *o++ = r.N;
        *o++ = r.dof;
        *o++ = r.mass;
        for (int c = 0; c < 3; c++) *o++ = r.I[c];
        for (int c = 0; c < 3; c++) *o++ = r.invI[c];
        for (int c = 0; c < 3; c++) *o++ = r.rcm[c];
        for (int c = 0; c < 3; c++) *o++ = r.pcm[c];
        for (int c = 0; c < 4; c++) *o++ = r.q[c];
        for (int c = 0; c < 4; c++) *o++ = r.pi[c];
        for (int c = 0; c < 3; c++) *o++ = r.force[c];
        for (int c = 0; c < 4; c++) *o++ = r.torque[c];
        *o++ = r.twoKt;
        *o++ = r.twoKr;
        *o++ = r.loc;
    });
}
