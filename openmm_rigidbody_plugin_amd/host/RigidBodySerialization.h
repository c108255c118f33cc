// XML serialization of RigidBodyIntegrator.
//
// Same document as the reference's serialization proxy
// (serialization/src/RigidBodyIntegratorProxy.cpp:40-67 with OpenMM's
// XmlSerializer): a root element carrying type="RigidBodyIntegrator",
// version="1", stepSize and constraintTolerance, and a <bodyIndices> child
// with one <bodyIndex index="..."/> per atom.  The rotation mode is written
// as one more attribute, rotationMode; a document without it (as the
// reference writes) reads back as mode 0, the reference's default.
// Doubles are written with 17 significant digits, so they read back exactly.
#pragma once

#include <iosfwd>
#include <string>

#include "RigidBodyIntegrator.h"

namespace RigidBodyPlugin {

class RigidBodyIntegratorProxy {
public:
    // writes `integrator` as an XML document whose root element is `rootName`
    static void serialize(const RigidBodyIntegrator& integrator, const std::string& rootName, std::ostream& out);
    // reads a document written by serialize (or by the reference's proxy);
    // throws rbhost::Exception on malformed input or an unsupported version
    static RigidBodyIntegrator* deserialize(std::istream& in);
};

}  // namespace RigidBodyPlugin
