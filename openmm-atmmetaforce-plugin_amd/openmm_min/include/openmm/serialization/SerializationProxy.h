// Minimal OpenMM-compatible core: converts objects of one class to and from
// SerializationNodes; a global registry maps C++ types and type names to proxies.
#pragma once
#include <string>
#include <typeinfo>
#include "openmm/serialization/SerializationNode.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT SerializationProxy {
public:
    explicit SerializationProxy(const std::string& typeName) : typeName(typeName) {}
    virtual ~SerializationProxy() {}
    const std::string& getTypeName() const { return typeName; }
    virtual void serialize(const void* object, SerializationNode& node) const = 0;
    virtual void* deserialize(const SerializationNode& node) const = 0;
    static void registerProxy(const std::type_info& type, const SerializationProxy* proxy);
    static const SerializationProxy& getProxy(const std::string& typeName);
    static const SerializationProxy& getProxy(const std::type_info& type);

private:
    std::string typeName;
};

}  // namespace OpenMM
