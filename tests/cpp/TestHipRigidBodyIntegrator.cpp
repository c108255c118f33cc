// C++ tests of the host plugin layer on the HIP platform, written as the
// reference's tests/TestRigidBodyIntegrator.h is: System + RigidBodyIntegrator
// + Context, checked against closed forms.  Exit status 0 = all passed.
//
//   testSingleBond       the reference's known-answer test
//                        (tests/TestRigidBodyIntegrator.h:49-83): two free
//                        atoms on a harmonic bond are an analytic oscillator
//   testRigidDimerSpin   a free rigid dimer keeps its bond length, its centre
//                        of mass moves uniformly, its angular momentum and
//                        kinetic energy are conserved, it rotates at |L|/I,
//                        and each atom moves at v_cm + w e_z x (r_i - c)
//   testHostForce        the force slot: a tether evaluated on the host
//                        (HostForce) drives the same trajectory as the
//                        device tether, alone and added to device forces
//   testApiErrors        the reference's exceptions (negative rotation mode,
//                        mode change after binding, body-index count, body
//                        constraints) and this platform's documented limits
//   testBookkeeping      RigidBodySystem counts, atom order and DOF
//   testSingleBondRefined  setComputeRefinedEnergies on testSingleBond's
//                        oscillator: refined kinetic energy + potential +
//                        potential refinement is the velocity-Verlet shadow
//                        energy, conserved far better than the plain energy
//   testVirtualSites     a rigid water with two massless tethered sites (an
//                        M site and a local-coordinates site)
//   testConstraints / testConstrainedClusters  the reference's constraint
//                        tests (tests/TestRigidBodyIntegrator.h:85-197,
//                        commented out there) with tethers standing in for
//                        NonbondedForce: distances held to 1e-4 / 2e-5,
//                        energy to 1% over 1000 steps, one DOF per constraint
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <tuple>
#include <vector>

#include "HipRigidBodyKernels.h"
#include "rigidbody_hip_plugin.h"

using namespace RigidBodyPlugin;
using namespace rbhost;

static int failures = 0;

// scaled comparisons: relative when |expected| > 1, absolute otherwise
static bool close(double expected, double found, double tol) {
    double scale = std::fabs(expected) > 1.0 ? std::fabs(expected) : 1.0;
    return std::fabs(expected - found) / scale <= tol;
}
#define CHECK_TOL(expected, found, tol)                                                                         \
    do {                                                                                                        \
        double e_ = (expected), f_ = (found);                                                                   \
        if (!close(e_, f_, (tol))) {                                                                            \
            std::printf("  %s:%d expected %.10g found %.10g (tol %g)\n", __FILE__, __LINE__, e_, f_, (double)(tol)); \
            throw Exception("check failed");                                                                    \
        }                                                                                                       \
    } while (0)
#define CHECK_VEC(expected, found, tol)                                                                 \
    do {                                                                                                \
        Vec3 e_ = (expected), f_ = (found);                                                             \
        double s_ = std::sqrt(e_.dot(e_));                                                              \
        s_ = s_ > 1.0 ? s_ : 1.0;                                                                       \
        Vec3 d_ = e_ - f_;                                                                              \
        if (!(std::sqrt(d_.dot(d_)) / s_ <= (tol))) {                                                   \
            std::printf("  %s:%d expected (%g %g %g) found (%g %g %g)\n", __FILE__, __LINE__, e_[0], e_[1], \
                        e_[2], f_[0], f_[1], f_[2]);                                                    \
            throw Exception("check failed");                                                            \
        }                                                                                               \
    } while (0)
#define CHECK_THROWS(stmt)                                                                    \
    do {                                                                                      \
        bool thrown_ = false;                                                                 \
        try {                                                                                 \
            stmt;                                                                             \
        } catch (const Exception&) {                                                          \
            thrown_ = true;                                                                   \
        }                                                                                     \
        if (!thrown_) {                                                                       \
            std::printf("  %s:%d expected an exception from: %s\n", __FILE__, __LINE__, #stmt); \
            throw Exception("check failed");                                                  \
        }                                                                                     \
    } while (0)

static std::map<std::string, std::string> props(const std::string& precision) {
    return {{"Precision", precision}};
}

void testSingleBond(Platform& platform, const std::string& precision) {
    System system;
    system.addParticle(2.0);
    system.addParticle(2.0);
    std::vector<int> bodyIndices(2, 0);
    RigidBodyIntegrator integrator(0.01, bodyIndices);
    HarmonicBondForce* forceField = new HarmonicBondForce();
    forceField->addBond(0, 1, 1.5, 1);
    system.addForce(forceField);
    Context context(system, integrator, platform, props(precision));
    std::vector<Vec3> positions(2);
    positions[0] = Vec3(-1, 0, 0);
    positions[1] = Vec3(1, 0, 0);
    context.setPositions(positions);

    const double freq = 1.0;
    State state = context.getState(State::Energy);
    const double initialEnergy = state.getKineticEnergy() + state.getPotentialEnergy();
    for (int i = 0; i < 1000; ++i) {
        state = context.getState(State::Positions | State::Velocities | State::Energy);
        double time = state.getTime();
        double expectedDist = 1.5 + 0.5 * std::cos(freq * time);
        CHECK_VEC(Vec3(-0.5 * expectedDist, 0, 0), state.getPositions()[0], 0.02);
        CHECK_VEC(Vec3(0.5 * expectedDist, 0, 0), state.getPositions()[1], 0.02);
        double expectedSpeed = -0.5 * freq * std::sin(freq * time);
        CHECK_VEC(Vec3(-0.5 * expectedSpeed, 0, 0), state.getVelocities()[0], 0.02);
        CHECK_VEC(Vec3(0.5 * expectedSpeed, 0, 0), state.getVelocities()[1], 0.02);
        double energy = state.getKineticEnergy() + state.getPotentialEnergy();
        CHECK_TOL(initialEnergy, energy, 0.01);
        integrator.step(1);
    }
    CHECK_TOL(10.0, context.getState(0).getTime(), 1e-5);
}

// a rigid dimer (masses m1, m2 at distance d) spinning freely about an axis
// perpendicular to the bond: |L| = I w with I = mu d^2, the bond turns by w t
void testRigidDimerSpin(Platform& platform, const std::string& precision, int mode) {
    System system;
    const double m1 = 3.0, m2 = 5.0, d = 0.4, w = 2.0;
    system.addParticle(m1);
    system.addParticle(m2);
    RigidBodyIntegrator integrator(0.002, std::vector<int>{7, 7});
    integrator.setRotationMode(mode);
    Context context(system, integrator, platform, props(precision));
    const double M = m1 + m2, r1 = d * m2 / M, r2 = d * m1 / M;
    const Vec3 vcm(0.3, -0.2, 0.1), c0(1.0, 2.0, 3.0);
    context.setPositions({c0 + Vec3(-r1, 0, 0), c0 + Vec3(r2, 0, 0)});
    // rotation about z at rate w: v_i = vcm + w e_z x (r_i - c)
    context.setVelocities({vcm + Vec3(0, -w * r1, 0), vcm + Vec3(0, w * r2, 0)});
    const double I = m1 * r1 * r1 + m2 * r2 * r2;
    const double ke0 = 0.5 * M * vcm.dot(vcm) + 0.5 * I * w * w;
    RigidBody b0 = integrator.getRigidBodySystem().getRigidBody(0);
    if (b0.dof != 5) throw Exception("a dimer is a linear body (dof 5)");
    CHECK_TOL(M, b0.mass, 1e-12);
    CHECK_TOL(I, b0.I[0], 1e-6);
    for (int s = 0; s < 500; s++) integrator.step(1);
    State st = context.getState(State::Positions | State::Velocities | State::Energy);
    const std::vector<Vec3>& R = st.getPositions();
    const std::vector<Vec3>& V = st.getVelocities();
    double t = st.getTime();
    Vec3 c = (R[0] * m1 + R[1] * m2) * (1.0 / M);
    CHECK_VEC(c0 + vcm * t, c, 1e-5);
    Vec3 bond = R[1] - R[0];
    CHECK_TOL(d, std::sqrt(bond.dot(bond)), 1e-6);
    CHECK_VEC(Vec3(d * std::cos(w * t), d * std::sin(w * t), 0), bond, 2e-4);
    Vec3 vc = (V[0] * m1 + V[1] * m2) * (1.0 / M);
    CHECK_VEC(vcm, vc, 1e-5);
    // each atom: v_i = v_cm + w e_z x (r_i - c) (RigidBody.cpp:159-168)
    for (int i = 0; i < 2; i++) CHECK_VEC(vcm + Vec3(0, 0, w).cross(R[i] - c), V[i], 2e-5);
    CHECK_TOL(ke0, st.getKineticEnergy(), 1e-5);
    std::vector<double> ke = integrator.getKineticEnergies();
    CHECK_TOL(0.5 * I * w * w, ke[1], 1e-5);
    std::vector<double> rk = integrator.getRefinedKineticEnergies();
    CHECK_TOL(ke[0], rk[0], 0.0);
    CHECK_TOL(ke[1], rk[1], 0.0);
    CHECK_TOL(0.0, integrator.getPotentialEnergyRefinement(), 0.0);
}

void testSingleBondRefined(Platform& platform) {
    System system;
    system.addParticle(2.0);
    system.addParticle(2.0);
    RigidBodyIntegrator integrator(0.01, std::vector<int>(2, 0));
    integrator.setComputeRefinedEnergies(true);
    HarmonicBondForce* forceField = new HarmonicBondForce();
    forceField->addBond(0, 1, 1.5, 1);
    system.addForce(forceField);
    Context context(system, integrator, platform, props("double"));
    context.setPositions({Vec3(-1, 0, 0), Vec3(1, 0, 0)});
    std::vector<double> rk = integrator.getRefinedKineticEnergies();
    CHECK_TOL(0.0, rk[0], 0.0);   // empty until the first step
    double plainMin = 1e30, plainMax = -1e30, shadowMin = 1e30, shadowMax = -1e30;
    for (int i = 0; i < 1000; ++i) {
        integrator.step(1);
        State state = context.getState(State::Energy);
        rk = integrator.getRefinedKineticEnergies();
        double plain = state.getKineticEnergy() + state.getPotentialEnergy();
        double shadow = rk[0] + rk[1] + state.getPotentialEnergy() + integrator.getPotentialEnergyRefinement();
        plainMin = std::min(plainMin, plain);
        plainMax = std::max(plainMax, plain);
        shadowMin = std::min(shadowMin, shadow);
        shadowMax = std::max(shadowMax, shadow);
    }
    if (!(plainMax - plainMin > 0.0 && shadowMax - shadowMin < (plainMax - plainMin) / 20)) {
        std::printf("  plain energy range %g, shadow energy range %g\n", plainMax - plainMin, shadowMax - shadowMin);
        throw Exception("the shadow energy is not conserved better than the plain energy");
    }
}

// deterministic uniform numbers in [-0.5, 0.5) (a 64-bit LCG)
struct Uniform {
    unsigned long long x;
    explicit Uniform(unsigned long long seed) : x(seed * 2862933555777941757ULL + 3037000493ULL) {}
    double operator()() {
        x = x * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(x >> 11) / 9007199254740992.0 - 0.5;
    }
};

static void runConstrained(Platform& platform, const std::string& precision, const std::vector<double>& masses,
                           const std::vector<Vec3>& positions, const std::vector<std::tuple<int, int, double>>& cons,
                           double distTol, unsigned long long seed) {
    const int n = (int)masses.size();
    System system;
    for (double m : masses) system.addParticle(m);
    for (const auto& c : cons) system.addConstraint(std::get<0>(c), std::get<1>(c), std::get<2>(c));
    Uniform u(seed);
    TetherForce* tether = new TetherForce();
    for (int i = 0; i < n; i++) tether->addParticle(i, 50.0, positions[i] + Vec3(0.4 * u(), 0.4 * u(), 0.4 * u()));
    system.addForce(tether);
    RigidBodyIntegrator integrator(0.001, std::vector<int>(n, 0));
    integrator.setConstraintTolerance(1e-5);
    Context context(system, integrator, platform, props(precision));
    context.setPositions(positions);
    std::vector<Vec3> v(n);
    for (int i = 0; i < n; i++) v[i] = Vec3(u(), u(), u());
    context.setVelocities(v);
    // RigidBodySystem.cpp:130: numFree - numConstraints (a free atom counts one)
    if (integrator.getRigidBodySystem().getNumDOF() != n - (int)cons.size()) throw Exception("numDOF");
    double e0 = 0.0;
    for (int i = 0; i < 1000; i++) {
        State st = context.getState(State::Positions | State::Energy);
        const std::vector<Vec3>& R = st.getPositions();
        for (const auto& c : cons) {
            Vec3 d = R[std::get<0>(c)] - R[std::get<1>(c)];
            CHECK_TOL(std::get<2>(c), std::sqrt(d.dot(d)), distTol);
        }
        double e = st.getPotentialEnergy() + st.getKineticEnergy();
        if (i == 1) e0 = e;
        else if (i > 1 && !(std::fabs(e - e0) <= 0.01 * std::fabs(e0)))
            throw Exception("energy drifted from " + std::to_string(e0) + " to " + std::to_string(e));
        integrator.step(1);
    }
}

void testConstraints(Platform& platform, const std::string& precision) {
    std::vector<double> masses;
    std::vector<Vec3> positions;
    for (int i = 0; i < 8; i++) {
        masses.push_back(i % 2 == 0 ? 5.0 : 10.0);
        positions.push_back(Vec3(i / 2, (i + 1) / 2, 0));
    }
    runConstrained(platform, precision, masses, positions,
                   {{0, 1, 1.0}, {1, 2, 1.0}, {2, 3, 1.0}, {4, 5, 1.0}, {6, 7, 1.0}}, 1e-4, 0);
}

void testConstrainedClusters(Platform& platform, const std::string& precision) {
    const double r2 = std::sqrt(2.0);
    runConstrained(platform, precision, {10.0, 10.0, 1.0, 1.0, 1.0, 1.0, 1.0},
                   {Vec3(0, 0, 0), Vec3(1, 0, 0), Vec3(-1, 0, 0), Vec3(0, 1, 0), Vec3(0, 0, 1), Vec3(2, 0, 0),
                    Vec3(1, 1, 0)},
                   {{0, 1, 1.0}, {0, 2, 1.0}, {0, 3, 1.0}, {0, 4, 1.0}, {1, 5, 1.0}, {1, 6, 1.0}, {2, 3, r2},
                    {2, 4, r2}, {3, 4, r2}, {5, 6, r2}},
                   2e-5, 1);
}

// a rigid water with a massless fourth site (TIP4P's M site as a
// three-particle average) that carries a tether: the site follows its
// definition, the energy is conserved, and sites are not actual atoms
void testVirtualSites(Platform& platform, const std::string& precision) {
    System system;
    system.addParticle(15.9994);
    system.addParticle(1.008);
    system.addParticle(1.008);
    system.addParticle(0.0);
    system.addParticle(0.0);
    const double a = 0.128, b = 0.128, c = 1.0 - a - b;
    system.setVirtualSite(3, new ThreeParticleAverageSite(0, 1, 2, c, a, b));
    // a second site in the water's own frame
    const Vec3 lp(0.01, 0.02, 0.015);
    system.setVirtualSite(4, new LocalCoordinatesSite(0, 1, 2, Vec3(1, 0, 0), Vec3(-1, 1, 0), Vec3(-1, 0, 1), lp));
    const double th = 104.52 * M_PI / 180.0, d = 0.09572;
    std::vector<Vec3> x = {Vec3(0, 0, 0), Vec3(d, 0, 0), Vec3(d * std::cos(th), d * std::sin(th), 0)};
    x.push_back(x[0] * c + x[1] * a + x[2] * b);
    auto local = [&](const std::vector<Vec3>& R) {
        Vec3 xd = R[1] - R[0], yd = R[2] - R[0];
        Vec3 ex = xd * (1.0 / std::sqrt(xd.dot(xd))), z = xd.cross(yd);
        Vec3 ez = z * (1.0 / std::sqrt(z.dot(z))), ey = ez.cross(ex);
        return R[0] + ex * lp[0] + ey * lp[1] + ez * lp[2];
    };
    x.push_back(local(x));
    TetherForce* tether = new TetherForce();
    for (int i = 0; i < 5; i++) tether->addParticle(i, 400.0, x[i] + Vec3(0.01 * (i + 1), -0.005 * i, 0.008));
    system.addForce(tether);
    RigidBodyIntegrator integrator(0.001, std::vector<int>{1, 1, 1, 0, 0});
    Context context(system, integrator, platform, props(precision));
    context.setPositions(x);
    context.setVelocities({Vec3(0.2, 0.1, 0.0), Vec3(-0.5, 0.8, 0.3), Vec3(0.4, -0.6, -0.2), Vec3(), Vec3()});
    const RigidBodySystem& bs = integrator.getRigidBodySystem();
    if (bs.getNumActualAtoms() != 3 || bs.getNumBodies() != 1 || bs.getNumFree() != 0) throw Exception("counts");
    double e0 = 0.0;
    for (int i = 0; i < 1000; i++) {
        State st = context.getState(State::Positions | State::Energy);
        const std::vector<Vec3>& R = st.getPositions();
        if (i > 0) {
            CHECK_VEC(R[0] * c + R[1] * a + R[2] * b, R[3], 2e-6);
            CHECK_VEC(local(R), R[4], 2e-6);
        }
        double e = st.getPotentialEnergy() + st.getKineticEnergy();
        if (i == 1) e0 = e;
        else if (i > 1 && !(std::fabs(e - e0) <= 0.01 * std::fabs(e0)))
            throw Exception("energy drifted from " + std::to_string(e0) + " to " + std::to_string(e));
        integrator.step(1);
    }
}

// tethered rigid waters and free atoms; the tether is split as asked
// between the device TetherForce (fraction kDevice of k) and a HostForce
// evaluating the rest on the host
static State runTethered(Platform& platform, const std::string& precision, int mode, double kDevice, int steps,
                         bool& multiStep) {
    const double k = 800.0;
    const Vec3 tmpl[3] = {Vec3(0, 0, 0.0066), Vec3(0, 0.0757, -0.0520), Vec3(0, -0.0757, -0.0520)};
    const double wm[3] = {15.9994, 1.008, 1.008};
    System system;
    std::vector<int> bodies;
    std::vector<Vec3> pos, vel;
    for (int b = 0; b < 27; b++) {
        Vec3 c(0.4 * (b % 3), 0.4 * ((b / 3) % 3), 0.4 * (b / 9));
        double a = 0.7 * b, ca = std::cos(a), sa = std::sin(a);
        for (int j = 0; j < 3; j++) {
            system.addParticle(wm[j]);
            bodies.push_back(b + 1);
            Vec3 t = tmpl[j];
            pos.push_back(c + Vec3(ca * t[0] - sa * t[1], sa * t[0] + ca * t[1], t[2]));
            vel.push_back(Vec3(0.3 * std::sin(1.3 * b + j), -0.2 * std::cos(0.7 * b - j), 0.25 * std::sin(b * j + 0.5)));
        }
    }
    for (int f = 0; f < 10; f++) {
        system.addParticle(39.948);
        bodies.push_back(0);
        pos.push_back(Vec3(1.3 + 0.1 * f, 0.2 * f, -0.3));
        vel.push_back(Vec3(0.1 * f - 0.4, 0.05 * f, 0.2));
    }
    const int n = system.getNumParticles();
    std::vector<Vec3> r0(n);
    for (int i = 0; i < n; i++) r0[i] = pos[i] + Vec3(0.01 * std::sin(i), -0.01 * std::cos(2.0 * i), 0.005);
    if (kDevice > 0.0) {
        TetherForce* tf = new TetherForce();
        for (int i = 0; i < n; i++) tf->addParticle(i, kDevice * k, r0[i]);
        system.addForce(tf);
    }
    if (kDevice < 1.0) {
        const double kh = (1.0 - kDevice) * k;
        system.addForce(new HostForce([r0, kh](const std::vector<Vec3>& x, std::vector<Vec3>& f) {
            double e = 0.0;
            for (size_t i = 0; i < x.size(); i++) {
                Vec3 d = x[i] - r0[i];
                f[i] = f[i] + d * (-kh);
                e += 0.5 * kh * d.dot(d);
            }
            return e;
        }));
    }
    RigidBodyIntegrator integrator(0.002, bodies);
    integrator.setRotationMode(mode);
    Context context(system, integrator, platform, props(precision));
    context.setPositions(pos);
    context.setVelocities(vel);
    // device forces only: one multi-step call; with a HostForce, one
    // execute() per step with the host force evaluation in between
    multiStep = kDevice >= 1.0;
    integrator.step(steps);
    return context.getState(State::Positions | State::Velocities | State::Energy);
}

void testHostForce(Platform& platform, const std::string& precision, int mode) {
    const double tol = precision == "mixed" ? 1e-6 : 1e-10;
    bool m0, m1, m2;
    State dev = runTethered(platform, precision, mode, 1.0, 20, m0);
    State host = runTethered(platform, precision, mode, 0.0, 20, m1);
    State split = runTethered(platform, precision, mode, 0.5, 20, m2);
    for (const State* s : {&host, &split}) {
        for (size_t i = 0; i < dev.getPositions().size(); i++) {
            CHECK_VEC(dev.getPositions()[i], s->getPositions()[i], tol);
            CHECK_VEC(dev.getVelocities()[i], s->getVelocities()[i], 100 * tol);
        }
        CHECK_TOL(dev.getKineticEnergy(), s->getKineticEnergy(), 100 * tol);
        CHECK_TOL(dev.getPotentialEnergy(), s->getPotentialEnergy(), 100 * tol);
        CHECK_TOL(dev.getTime(), s->getTime(), 1e-12);
    }
}

void testApiErrors(Platform& platform) {
    RigidBodyIntegrator a(0.001, std::vector<int>{1, 1});
    CHECK_THROWS(a.setRotationMode(-1));
    {
        System system;
        system.addParticle(1.0);
        system.addParticle(1.0);
        RigidBodyIntegrator integrator(0.001, std::vector<int>{1, 1});
        Context context(system, integrator, platform);
        CHECK_THROWS(integrator.setRotationMode(3));
        CHECK_THROWS(integrator.setComputeRefinedEnergies(true));
        RigidBodyIntegrator other(0.001, std::vector<int>{1, 1});
        CHECK_THROWS(other.step(1));
    }
    {
        System system;
        system.addParticle(1.0);
        system.addParticle(1.0);
        RigidBodyIntegrator integrator(0.001, std::vector<int>{1, 1, 0});
        CHECK_THROWS(Context(system, integrator, platform));
    }
    {
        System system;
        for (int i = 0; i < 3; i++) system.addParticle(1.0);
        system.addConstraint(0, 2, 1.0);
        RigidBodyIntegrator integrator(0.001, std::vector<int>{1, 1, 0});
        CHECK_THROWS(Context(system, integrator, platform));
    }
    {
        // this platform's limits: single-atom bodies, massless particles in a body
        System system;
        system.addParticle(1.0);
        system.addParticle(1.0);
        RigidBodyIntegrator integrator(0.001, std::vector<int>{1, 2});
        CHECK_THROWS(Context(system, integrator, platform));
        System massless;
        massless.addParticle(1.0);
        massless.addParticle(1.0);
        massless.addParticle(0.0);
        RigidBodyIntegrator integrator2(0.001, std::vector<int>{1, 1, 1});
        CHECK_THROWS(Context(massless, integrator2, platform));
        // a constraint cannot involve a massless particle (testConstrainedMasslessParticles)
        System cm;
        cm.addParticle(0.0);
        cm.addParticle(1.0);
        cm.addConstraint(0, 1, 1.5);
        RigidBodyIntegrator integrator3(0.01, std::vector<int>{0, 0});
        CHECK_THROWS(Context(cm, integrator3, platform));
    }
    {
        System system;
        system.addParticle(1.0);
        RigidBodyIntegrator integrator(0.001, std::vector<int>{0});
        CHECK_THROWS(Context(system, integrator, platform, {{"Precision", "half"}}));
    }
    {
        // the integrator is free again once its context is gone
        System system;
        system.addParticle(1.0);
        system.addParticle(1.0);
        RigidBodyIntegrator integrator(0.001, std::vector<int>{4, 4});
        { Context context(system, integrator, platform); }
        integrator.setRotationMode(2);
        Context context(system, integrator, platform);
        if (integrator.getRotationMode() != 2) throw Exception("rotation mode not kept");
    }
}

void testBookkeeping(Platform& platform) {
    // labels with gaps and out of order; free atoms interleaved
    std::vector<int> labels = {9, 0, 4, 9, -1, 4, 4, 0, 9, 12, 12};
    System system;
    for (size_t i = 0; i < labels.size(); i++) system.addParticle(1.0 + i);
    RigidBodyIntegrator integrator(0.001, labels);
    integrator.setRotationMode(5);
    Context context(system, integrator, platform);
    std::vector<Vec3> pos;
    for (size_t i = 0; i < labels.size(); i++) pos.push_back(Vec3(0.3 * i, 0.1 * (i % 3), 0.05 * (i * i % 5)));
    context.setPositions(pos);
    const RigidBodySystem& bs = integrator.getRigidBodySystem();
    if (bs.getNumFree() != 3 || bs.getNumBodies() != 3 || bs.getNumActualAtoms() != 11 || bs.getNumBodyAtoms() != 8)
        throw Exception("counts");
    // free atoms ascending, then body 1 (label 4), body 2 (label 9), body 3 (label 12)
    const int expected[] = {1, 4, 7, 2, 5, 6, 0, 3, 8, 9, 10};
    for (int i = 0; i < 11; i++)
        if (bs.getAtomIndex(i) != expected[i]) throw Exception("atom index");
    int dof = 3;
    for (int b = 0; b < 3; b++) dof += bs.getRigidBody(b).dof;
    if (bs.getNumDOF() != dof) throw Exception("numDOF");
    if (bs.getRigidBody(2).dof != 5 || bs.getRigidBody(0).dof != 6) throw Exception("body dof");
}

int main() {
    registerRigidBodyHipKernelFactories();
    Platform& platform = Platform::getPlatformByName("HIP");
    std::vector<std::pair<std::string, std::function<void()>>> tests = {
        {"testSingleBond mixed", [&] { testSingleBond(platform, "mixed"); }},
        {"testSingleBond double", [&] { testSingleBond(platform, "double"); }},
        {"testRigidDimerSpin mixed NO-SQUISH", [&] { testRigidDimerSpin(platform, "mixed", 4); }},
        {"testRigidDimerSpin double exact", [&] { testRigidDimerSpin(platform, "double", 0); }},
        {"testApiErrors", [&] { testApiErrors(platform); }},
        {"testBookkeeping", [&] { testBookkeeping(platform); }},
        {"testSingleBondRefined", [&] { testSingleBondRefined(platform); }},
        {"testConstraints mixed", [&] { testConstraints(platform, "mixed"); }},
        {"testConstraints double", [&] { testConstraints(platform, "double"); }},
        {"testConstrainedClusters mixed", [&] { testConstrainedClusters(platform, "mixed"); }},
        {"testConstrainedClusters double", [&] { testConstrainedClusters(platform, "double"); }},
        {"testVirtualSites mixed", [&] { testVirtualSites(platform, "mixed"); }},
        {"testVirtualSites double", [&] { testVirtualSites(platform, "double"); }},
        {"testHostForce mixed NO-SQUISH", [&] { testHostForce(platform, "mixed", 10); }},
        {"testHostForce mixed exact", [&] { testHostForce(platform, "mixed", 0); }},
        {"testHostForce double exact", [&] { testHostForce(platform, "double", 0); }},
    };
    for (auto& t : tests) {
        std::printf("%s ...\n", t.first.c_str());
        std::fflush(stdout);
        try {
            t.second();
            std::printf("  ok\n");
        } catch (const std::exception& e) {
            std::printf("  FAILED: %s\n", e.what());
            failures++;
        }
        std::fflush(stdout);
    }
    std::printf(failures ? "FAIL (%d)\n" : "Done\n", failures);
    return failures ? 1 : 0;
}
