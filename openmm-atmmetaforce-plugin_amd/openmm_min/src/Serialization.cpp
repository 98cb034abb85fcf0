// Minimal OpenMM-compatible core: SerializationNode, the proxy registry and
// the XML writer/reader used by XmlSerializer.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <istream>
#include <map>
#include <ostream>
#include <sstream>
#include <typeindex>

#include "openmm/OpenMMException.h"
#include "openmm/Platform.h"
#include "openmm/State.h"
#include "openmm/serialization/SerializationNode.h"
#include "openmm/serialization/SerializationProxy.h"
#include "openmm/serialization/XmlSerializer.h"

namespace OpenMM {

// ------------------------------------------------------------------ SerializationNode

const SerializationNode& SerializationNode::getChildNode(const std::string& n) const {
    for (const SerializationNode& c : children)
        if (c.name == n) return c;
    throw OpenMMException("Unknown child '" + n + "' in node '" + name + "'");
}

SerializationNode& SerializationNode::getChildNode(const std::string& n) {
    for (SerializationNode& c : children)
        if (c.name == n) return c;
    throw OpenMMException("Unknown child '" + n + "' in node '" + name + "'");
}

SerializationNode& SerializationNode::createChildNode(const std::string& n) {
    children.emplace_back();
    children.back().name = n;
    return children.back();
}

const std::string& SerializationNode::getStringProperty(const std::string& n) const {
    auto it = properties.find(n);
    if (it == properties.end()) throw OpenMMException("Unknown property '" + n + "' in node '" + name + "'");
    return it->second;
}

const std::string& SerializationNode::getStringProperty(const std::string& n, const std::string& def) const {
    auto it = properties.find(n);
    return it == properties.end() ? def : it->second;
}

SerializationNode& SerializationNode::setStringProperty(const std::string& n, const std::string& v) {
    properties[n] = v;
    return *this;
}

int SerializationNode::getIntProperty(const std::string& n) const {
    const std::string& s = getStringProperty(n);
    char* end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end) throw OpenMMException("Property '" + n + "' is not an integer: " + s);
    return (int)v;
}

int SerializationNode::getIntProperty(const std::string& n, int def) const {
    return hasProperty(n) ? getIntProperty(n) : def;
}

SerializationNode& SerializationNode::setIntProperty(const std::string& n, int v) {
    properties[n] = std::to_string(v);
    return *this;
}

bool SerializationNode::getBoolProperty(const std::string& n) const {
    const std::string& s = getStringProperty(n);
    if (s == "1" || s == "true") return true;
    if (s == "0" || s == "false") return false;
    throw OpenMMException("Property '" + n + "' is not a boolean: " + s);
}

bool SerializationNode::getBoolProperty(const std::string& n, bool def) const {
    return hasProperty(n) ? getBoolProperty(n) : def;
}

SerializationNode& SerializationNode::setBoolProperty(const std::string& n, bool v) {
    properties[n] = v ? "1" : "0";
    return *this;
}

double SerializationNode::getDoubleProperty(const std::string& n) const {
    const std::string& s = getStringProperty(n);
    char* end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end) throw OpenMMException("Property '" + n + "' is not a number: " + s);
    return v;
}

double SerializationNode::getDoubleProperty(const std::string& n, double def) const {
    return hasProperty(n) ? getDoubleProperty(n) : def;
}

// 17 significant digits: every double survives a write/read round trip.
SerializationNode& SerializationNode::setDoubleProperty(const std::string& n, double v) {
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    properties[n] = buf;
    return *this;
}

// ------------------------------------------------------------------ proxies

static std::map<std::string, const SerializationProxy*>& byName() {
    static std::map<std::string, const SerializationProxy*> m;
    return m;
}

static std::map<std::type_index, const SerializationProxy*>& byType() {
    static std::map<std::type_index, const SerializationProxy*> m;
    return m;
}

void SerializationProxy::registerProxy(const std::type_info& type, const SerializationProxy* proxy) {
    byName()[proxy->getTypeName()] = proxy;
    byType()[std::type_index(type)] = proxy;
}

const SerializationProxy& SerializationProxy::getProxy(const std::string& typeName) {
    auto it = byName().find(typeName);
    if (it == byName().end()) throw OpenMMException("There is no serialization proxy registered for type " + typeName);
    return *it->second;
}

const SerializationProxy& SerializationProxy::getProxy(const std::type_info& type) {
    auto it = byType().find(std::type_index(type));
    if (it == byType().end())
        throw OpenMMException(std::string("There is no serialization proxy registered for type ") + type.name());
    return *it->second;
}

// ------------------------------------------------------------------ XML

static void escape(std::ostream& o, const std::string& s) {
    for (char ch : s) {
        switch (ch) {
            case '&': o << "&amp;"; break;
            case '<': o << "&lt;"; break;
            case '>': o << "&gt;"; break;
            case '"': o << "&quot;"; break;
            case '\'': o << "&apos;"; break;
            default: o << ch;
        }
    }
}

static void writeNode(const SerializationNode& node, std::ostream& o, int depth) {
    for (int i = 0; i < depth; ++i) o << '\t';
    o << '<' << node.getName();
    for (auto& kv : node.getProperties()) {
        o << ' ' << kv.first << "=\"";
        escape(o, kv.second);
        o << '"';
    }
    if (node.getChildren().empty()) {
        o << "/>\n";
        return;
    }
    o << ">\n";
    for (const SerializationNode& c : node.getChildren()) writeNode(c, o, depth + 1);
    for (int i = 0; i < depth; ++i) o << '\t';
    o << "</" << node.getName() << ">\n";
}

void XmlSerializer::writeXml(const SerializationNode& node, std::ostream& stream) {
    stream << "<?xml version=\"1.0\" ?>\n";
    writeNode(node, stream, 0);
}

namespace {

struct Reader {
    std::string s;
    size_t p = 0;
    [[noreturn]] void error(const std::string& what) {
        throw OpenMMException("XML parse error at offset " + std::to_string(p) + ": " + what);
    }
    void skipSpace() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
    }
    bool startsWith(const char* t) const { return s.compare(p, std::char_traits<char>::length(t), t) == 0; }
    // skip declarations, comments and processing instructions
    void skipMisc() {
        for (;;) {
            skipSpace();
            if (startsWith("<?")) {
                size_t e = s.find("?>", p);
                if (e == std::string::npos) error("unterminated declaration");
                p = e + 2;
            } else if (startsWith("<!--")) {
                size_t e = s.find("-->", p);
                if (e == std::string::npos) error("unterminated comment");
                p = e + 3;
            } else if (startsWith("<!")) {
                size_t e = s.find('>', p);
                if (e == std::string::npos) error("unterminated declaration");
                p = e + 1;
            } else {
                return;
            }
        }
    }
    std::string name() {
        size_t b = p;
        while (p < s.size() && (std::isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '-' || s[p] == '.' || s[p] == ':')) ++p;
        if (p == b) error("expected a name");
        return s.substr(b, p - b);
    }
    std::string unescape(const std::string& v) {
        std::string o;
        for (size_t i = 0; i < v.size(); ++i) {
            if (v[i] != '&') { o += v[i]; continue; }
            size_t e = v.find(';', i);
            if (e == std::string::npos) error("bad entity");
            std::string ent = v.substr(i + 1, e - i - 1);
            if (ent == "amp") o += '&';
            else if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else if (!ent.empty() && ent[0] == '#') o += (char)std::strtol(ent.c_str() + (ent[1] == 'x' ? 2 : 1), nullptr, ent[1] == 'x' ? 16 : 10);
            else error("unknown entity &" + ent + ";");
            i = e;
        }
        return o;
    }
    void element(SerializationNode& node) {
        if (p >= s.size() || s[p] != '<') error("expected '<'");
        ++p;
        node.setName(name());
        for (;;) {
            skipSpace();
            if (p >= s.size()) error("unterminated element");
            if (s[p] == '/') {
                if (p + 1 >= s.size() || s[p + 1] != '>') error("expected '/>'");
                p += 2;
                return;
            }
            if (s[p] == '>') {
                ++p;
                break;
            }
            std::string key = name();
            skipSpace();
            if (p >= s.size() || s[p] != '=') error("expected '='");
            ++p;
            skipSpace();
            if (p >= s.size() || (s[p] != '"' && s[p] != '\'')) error("expected a quoted value");
            char q = s[p++];
            size_t e = s.find(q, p);
            if (e == std::string::npos) error("unterminated value");
            node.setStringProperty(key, unescape(s.substr(p, e - p)));
            p = e + 1;
        }
        for (;;) {
            skipMisc();
            if (p >= s.size()) error("missing end tag for " + node.getName());
            if (startsWith("</")) {
                p += 2;
                std::string n = name();
                if (n != node.getName()) error("mismatched end tag " + n);
                skipSpace();
                if (p >= s.size() || s[p] != '>') error("expected '>'");
                ++p;
                return;
            }
            if (s[p] == '<') {
                node.getChildren().emplace_back();
                element(node.getChildren().back());
            } else {
                while (p < s.size() && s[p] != '<') ++p;  // character data is ignored
            }
        }
    }
};

}  // namespace

// ------------------------------------------------------------------ State

// State checkpoints in OpenMM's State XML layout (version 1): root attributes
// openmmVersion and time; children PeriodicBoxVectors (A, B, C), Parameters
// (one attribute per global parameter, e.g. the ATM parameters), Energies,
// Positions/Velocities/Forces (one x, y, z element per particle) and
// IntegratorParameters.  Reads files such as the ATM examples' equilibrated
// checkpoints (example/abfe/temoa-g1-equil.xml).
class StateProxy : public SerializationProxy {
public:
    StateProxy() : SerializationProxy("State") {}
    void serialize(const void* object, SerializationNode& node) const override {
        const State& s = *reinterpret_cast<const State*>(object);
        node.setIntProperty("version", 1);
        node.setStringProperty("openmmVersion", Platform::getOpenMMVersion());
        node.setDoubleProperty("time", s.time);
        SerializationNode& box = node.createChildNode("PeriodicBoxVectors");
        const char* names[3] = {"A", "B", "C"};
        for (int k = 0; k < 3; ++k)
            box.createChildNode(names[k]).setDoubleProperty("x", s.box[k][0]).setDoubleProperty("y", s.box[k][1])
                .setDoubleProperty("z", s.box[k][2]);
        if (s.types & State::Parameters) {
            SerializationNode& pn = node.createChildNode("Parameters");
            for (auto& kv : s.parameters) pn.setDoubleProperty(kv.first, kv.second);
        }
        if (s.types & State::Energy)
            node.createChildNode("Energies").setDoubleProperty("KineticEnergy", s.kineticEnergy)
                .setDoubleProperty("PotentialEnergy", s.potentialEnergy);
        auto list = [&](const char* outer, const char* inner, const std::vector<Vec3>& v) {
            SerializationNode& n = node.createChildNode(outer);
            for (const Vec3& x : v)
                n.createChildNode(inner).setDoubleProperty("x", x[0]).setDoubleProperty("y", x[1]).setDoubleProperty("z", x[2]);
        };
        if (s.types & State::Positions) list("Positions", "Position", s.positions);
        if (s.types & State::Velocities) list("Velocities", "Velocity", s.velocities);
        if (s.types & State::Forces) list("Forces", "Force", s.forces);
        node.createChildNode("IntegratorParameters");
    }
    void* deserialize(const SerializationNode& node) const override {
        if (node.getIntProperty("version") != 1) throw OpenMMException("Unsupported version number");
        State* s = new State();
        try {
            s->time = node.getDoubleProperty("time");
            int types = 0;
            for (const SerializationNode& c : node.getChildren()) {
                const std::string& n = c.getName();
                auto vecs = [&](std::vector<Vec3>& v) {
                    v.clear();
                    for (const SerializationNode& e : c.getChildren())
                        v.emplace_back(e.getDoubleProperty("x"), e.getDoubleProperty("y"), e.getDoubleProperty("z"));
                };
                if (n == "PeriodicBoxVectors") {
                    const char* names[3] = {"A", "B", "C"};
                    for (int k = 0; k < 3; ++k) {
                        const SerializationNode& b = c.getChildNode(names[k]);
                        s->box[k] = Vec3(b.getDoubleProperty("x"), b.getDoubleProperty("y"), b.getDoubleProperty("z"));
                    }
                } else if (n == "Parameters") {
                    for (auto& kv : c.getProperties()) s->parameters[kv.first] = c.getDoubleProperty(kv.first);
                    types |= State::Parameters;
                } else if (n == "Energies") {
                    s->kineticEnergy = c.getDoubleProperty("KineticEnergy");
                    s->potentialEnergy = c.getDoubleProperty("PotentialEnergy");
                    types |= State::Energy;
                } else if (n == "Positions") {
                    vecs(s->positions);
                    types |= State::Positions;
                } else if (n == "Velocities") {
                    vecs(s->velocities);
                    types |= State::Velocities;
                } else if (n == "Forces") {
                    vecs(s->forces);
                    types |= State::Forces;
                }
            }
            s->types = types;
        } catch (...) {
            delete s;
            throw;
        }
        return s;
    }
};

namespace {
struct RegisterCoreProxies {
    RegisterCoreProxies() {
        static StateProxy state;
        SerializationProxy::registerProxy(typeid(State), &state);
    }
} registerCoreProxies;
}  // namespace

void XmlSerializer::readXml(std::istream& stream, SerializationNode& node) {
    Reader r;
    std::stringstream b;
    b << stream.rdbuf();
    r.s = b.str();
    r.skipMisc();
    r.element(node);
}

}  // namespace OpenMM
