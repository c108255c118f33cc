// Serialization of RigidBodyIntegrator, as the reference's
// serialization/tests/TestSerializeRigidBodyIntegrator.cpp checks it (round
// trip of step size, constraint tolerance and body indices), plus the rotation
// mode, exact doubles, a document in the reference's own form (no
// rotationMode attribute) and the errors.  Host code only: no device needed.
// Exit status 0 = all passed.
#include <cstdio>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "RigidBodySerialization.h"

using namespace RigidBodyPlugin;
using rbhost::Exception;

#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("  %s:%d check failed: %s\n", __FILE__, __LINE__, #cond); \
            throw Exception("check failed");                                     \
        }                                                                        \
    } while (0)

void testSerializeRigidBodyIntegrator() {
    std::vector<int> bodyIndices = {5, 4, 3, 2, 1};
    RigidBodyIntegrator intg(0.00342, bodyIndices);
    std::stringstream ss;
    RigidBodyIntegratorProxy::serialize(intg, "RigidBodyIntegrator", ss);
    RigidBodyIntegrator* intg2 = RigidBodyIntegratorProxy::deserialize(ss);
    CHECK(intg.getConstraintTolerance() == intg2->getConstraintTolerance());
    CHECK(intg.getStepSize() == intg2->getStepSize());
    CHECK(intg2->getBodyIndices() == bodyIndices);
    delete intg2;
}

void testRoundTripDetails() {
    std::vector<int> bodyIndices = {0, -3, 7, 7, 2147483647, 12, 0};
    RigidBodyIntegrator intg(0.1 + 0.2, bodyIndices);      // a double with no short decimal form
    intg.setConstraintTolerance(1.0 / 3.0);
    intg.setRotationMode(17);
    std::stringstream ss;
    RigidBodyIntegratorProxy::serialize(intg, "Integrator", ss);
    RigidBodyIntegrator* back = RigidBodyIntegratorProxy::deserialize(ss);
    CHECK(back->getStepSize() == 0.1 + 0.2);
    CHECK(back->getConstraintTolerance() == 1.0 / 3.0);
    CHECK(back->getRotationMode() == 17);
    CHECK(back->getBodyIndices() == bodyIndices);
    delete back;
}

void testReferenceDocument() {
    // the form the reference writes: no rotationMode, children, a comment
    std::string doc =
        "<?xml version=\"1.0\" ?>\n<!-- written elsewhere -->\n"
        "<RigidBodyIntegrator type=\"RigidBodyIntegrator\" version=\"1\" stepSize=\".002\" "
        "constraintTolerance=\"1e-05\">\n  <bodyIndices>\n    <bodyIndex index=\"1\"/>\n"
        "    <bodyIndex index='1' />\n    <bodyIndex index=\"0\"></bodyIndex>\n  </bodyIndices>\n"
        "</RigidBodyIntegrator>\n";
    std::stringstream ss(doc);
    RigidBodyIntegrator* intg = RigidBodyIntegratorProxy::deserialize(ss);
    CHECK(intg->getStepSize() == 0.002);
    CHECK(intg->getConstraintTolerance() == 1e-5);
    CHECK(intg->getRotationMode() == 0);
    CHECK((intg->getBodyIndices() == std::vector<int>{1, 1, 0}));
    delete intg;
}

void testErrors() {
    const char* bad[] = {
        "",                                                                         // empty
        "<RigidBodyIntegrator version=\"2\" stepSize=\"1\" constraintTolerance=\"1\"><bodyIndices/></RigidBodyIntegrator>",
        "<RigidBodyIntegrator version=\"1\" constraintTolerance=\"1\"><bodyIndices/></RigidBodyIntegrator>",
        "<RigidBodyIntegrator version=\"1\" stepSize=\"x\" constraintTolerance=\"1\"><bodyIndices/></RigidBodyIntegrator>",
        "<RigidBodyIntegrator version=\"1\" stepSize=\"1\" constraintTolerance=\"1\"></RigidBodyIntegrator>",
        "<RigidBodyIntegrator version=\"1\" stepSize=\"1\" constraintTolerance=\"1\"><bodyIndices></RigidBodyIntegrator>",
        "<Other type=\"VerletIntegrator\" version=\"1\" stepSize=\"1\" constraintTolerance=\"1\"><bodyIndices/></Other>",
        "<RigidBodyIntegrator version=\"1\" stepSize=\"1\" constraintTolerance=\"1\" rotationMode=\"-1\"><bodyIndices/></RigidBodyIntegrator>",
    };
    for (const char* doc : bad) {
        std::stringstream ss(doc);
        bool thrown = false;
        try {
            delete RigidBodyIntegratorProxy::deserialize(ss);
        } catch (const Exception&) {
            thrown = true;
        }
        if (!thrown) std::printf("  accepted: %s\n", doc);
        CHECK(thrown);
    }
}

int main() {
    std::vector<std::pair<std::string, std::function<void()>>> tests = {
        {"testSerializeRigidBodyIntegrator", testSerializeRigidBodyIntegrator},
        {"testRoundTripDetails", testRoundTripDetails},
        {"testReferenceDocument", testReferenceDocument},
        {"testErrors", testErrors},
    };
    int failures = 0;
    for (auto& t : tests) {
        std::printf("%s ...\n", t.first.c_str());
        try {
            t.second();
            std::printf("  ok\n");
        } catch (const std::exception& e) {
            std::printf("  FAILED: %s\n", e.what());
            failures++;
        }
    }
    std::printf(failures ? "FAIL (%d)\n" : "Done\n", failures);
    return failures ? 1 : 0;
}
