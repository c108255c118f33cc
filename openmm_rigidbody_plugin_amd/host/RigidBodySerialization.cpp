// XML serialization of RigidBodyIntegrator (see RigidBodySerialization.h).
#include "RigidBodySerialization.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <istream>
#include <iterator>
#include <ostream>
#include <utility>
#include <vector>

namespace RigidBodyPlugin {

using rbhost::Exception;

namespace {

struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> children;

    const std::string* attr(const std::string& key) const {
        for (const auto& kv : attrs)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    const XmlNode& child(const std::string& key) const {
        for (const XmlNode& c : children)
            if (c.name == key) return c;
        throw Exception("Missing XML element <" + key + "> in <" + name + ">");
    }
};

// a reader for the element/attribute subset of XML these documents use:
// declarations, comments, nested and self-closing elements, quoted
// attributes with the five predefined entities; text content is ignored
class XmlReader {
public:
    explicit XmlReader(std::string text) : s(std::move(text)) {}

    XmlNode document() {
        skipMisc();
        XmlNode root = element();
        skipMisc();
        if (pos != s.size()) fail("content after the root element");
        return root;
    }

private:
    std::string s;
    size_t pos = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw Exception("Malformed XML at offset " + std::to_string(pos) + ": " + what);
    }
    bool starts(const char* p) const { return s.compare(pos, std::char_traits<char>::length(p), p) == 0; }
    void skipSpace() {
        while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\n' || s[pos] == '\r')) pos++;
    }
    void skipPast(const char* end) {
        size_t e = s.find(end, pos);
        if (e == std::string::npos) fail(std::string("missing ") + end);
        pos = e + std::char_traits<char>::length(end);
    }
    // whitespace, <?...?> and <!--...-->
    void skipMisc() {
        for (;;) {
            skipSpace();
            if (starts("<?")) skipPast("?>");
            else if (starts("<!--")) skipPast("-->");
            else return;
        }
    }
    std::string name() {
        size_t b = pos;
        while (pos < s.size() && (std::isalnum((unsigned char)s[pos]) || s[pos] == '_' || s[pos] == '-' ||
                                  s[pos] == ':' || s[pos] == '.'))
            pos++;
        if (pos == b) fail("expected a name");
        return s.substr(b, pos - b);
    }
    std::string unescape(const std::string& v) const {
        std::string out;
        for (size_t i = 0; i < v.size(); i++) {
            if (v[i] != '&') {
                out += v[i];
                continue;
            }
            size_t e = v.find(';', i);
            if (e == std::string::npos) throw Exception("Malformed XML entity in \"" + v + "\"");
            std::string ent = v.substr(i + 1, e - i - 1);
            if (ent == "quot") out += '"';
            else if (ent == "amp") out += '&';
            else if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "apos") out += '\'';
            else throw Exception("Unknown XML entity &" + ent + ";");
            i = e;
        }
        return out;
    }
    XmlNode element() {
        if (pos >= s.size() || s[pos] != '<') fail("expected '<'");
        pos++;
        XmlNode node;
        node.name = name();
        for (;;) {
            skipSpace();
            if (starts("/>")) {
                pos += 2;
                return node;
            }
            if (starts(">")) {
                pos++;
                break;
            }
            std::string key = name();
            skipSpace();
            if (pos >= s.size() || s[pos] != '=') fail("expected '=' after attribute " + key);
            pos++;
            skipSpace();
            if (pos >= s.size() || (s[pos] != '"' && s[pos] != '\'')) fail("expected a quoted value");
            char q = s[pos++];
            size_t e = s.find(q, pos);
            if (e == std::string::npos) fail("unterminated attribute value");
            node.attrs.emplace_back(key, unescape(s.substr(pos, e - pos)));
            pos = e + 1;
        }
        for (;;) {
            // text between elements is not part of these documents' data
            while (pos < s.size() && s[pos] != '<') pos++;
            if (starts("<!--")) {
                skipPast("-->");
                continue;
            }
            if (starts("</")) {
                pos += 2;
                std::string closing = name();
                if (closing != node.name) fail("</" + closing + "> closes <" + node.name + ">");
                skipSpace();
                if (pos >= s.size() || s[pos] != '>') fail("expected '>'");
                pos++;
                return node;
            }
            if (pos >= s.size()) fail("unterminated element <" + node.name + ">");
            node.children.push_back(element());
        }
    }
};

std::string fmt(double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

double to_double(const XmlNode& n, const std::string& key) {
    const std::string* v = n.attr(key);
    if (!v) throw Exception("Missing property " + key);
    char* end = nullptr;
    double d = std::strtod(v->c_str(), &end);
    if (end == v->c_str() || *end != '\0') throw Exception("Property " + key + " is not a number: " + *v);
    return d;
}

long to_long(const XmlNode& n, const std::string& key) {
    const std::string* v = n.attr(key);
    if (!v) throw Exception("Missing property " + key);
    char* end = nullptr;
    long d = std::strtol(v->c_str(), &end, 10);
    if (end == v->c_str() || *end != '\0') throw Exception("Property " + key + " is not an integer: " + *v);
    return d;
}

}  // namespace

void RigidBodyIntegratorProxy::serialize(const RigidBodyIntegrator& integrator, const std::string& rootName,
                                         std::ostream& out) {
    out << "<?xml version=\"1.0\" ?>\n";
    out << "<" << rootName << " type=\"RigidBodyIntegrator\" version=\"1\" stepSize=\"" << fmt(integrator.getStepSize())
        << "\" constraintTolerance=\"" << fmt(integrator.getConstraintTolerance()) << "\" rotationMode=\""
        << integrator.getRotationMode() << "\">\n";
    out << "\t<bodyIndices>\n";
    for (int index : integrator.getBodyIndices()) out << "\t\t<bodyIndex index=\"" << index << "\"/>\n";
    out << "\t</bodyIndices>\n";
    out << "</" << rootName << ">\n";
}

RigidBodyIntegrator* RigidBodyIntegratorProxy::deserialize(std::istream& in) {
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    XmlNode root = XmlReader(std::move(text)).document();
    const std::string* type = root.attr("type");
    if (type && *type != "RigidBodyIntegrator")
        throw Exception("The document holds a " + *type + ", not a RigidBodyIntegrator");
    if (to_long(root, "version") != 1) throw Exception("Unsupported version number");
    double stepSize = to_double(root, "stepSize");
    double tolerance = to_double(root, "constraintTolerance");
    std::vector<int> bodyIndices;
    for (const XmlNode& c : root.child("bodyIndices").children) {
        if (c.name != "bodyIndex") continue;
        long index = to_long(c, "index");
        if (index < -2147483647L - 1 || index > 2147483647L) throw Exception("Body index out of range");
        bodyIndices.push_back((int)index);
    }
    RigidBodyIntegrator* integrator = new RigidBodyIntegrator(stepSize, bodyIndices);
    integrator->setConstraintTolerance(tolerance);
    if (root.attr("rotationMode")) {
        try {
            integrator->setRotationMode((int)to_long(root, "rotationMode"));
        } catch (...) {
            delete integrator;
            throw;
        }
    }
    return integrator;
}

}  // namespace RigidBodyPlugin
