// Minimal OpenMM-compatible core: XML (de)serialization through the registered
// proxies.  Layout: an XML declaration, then one element per node, attributes
// in name order, children indented by one tab per level.
#pragma once
#include <iosfwd>
#include <sstream>
#include <string>
#include <typeinfo>
#include "openmm/serialization/SerializationNode.h"
#include "openmm/serialization/SerializationProxy.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT XmlSerializer {
public:
    template <class T> static void serialize(const T* object, const std::string& rootName, std::ostream& stream) {
        const SerializationProxy& proxy = SerializationProxy::getProxy(typeid(*object));
        SerializationNode node;
        node.setName(rootName);
        proxy.serialize(object, node);
        if (node.hasProperty("type")) throw OpenMMException("The SerializationProxy tried to set the reserved property 'type'");
        node.setStringProperty("type", proxy.getTypeName());
        writeXml(node, stream);
    }
    template <class T> static T* deserialize(std::istream& stream) {
        SerializationNode node;
        readXml(stream, node);
        const SerializationProxy& proxy = SerializationProxy::getProxy(node.getStringProperty("type"));
        void* obj = proxy.deserialize(node);
        T* t = reinterpret_cast<T*>(obj);
        return t;
    }
    template <class T> static T* clone(const T& object) {
        std::stringstream s;
        serialize<T>(&object, "Clone", s);
        return deserialize<T>(s);
    }
    static void writeXml(const SerializationNode& node, std::ostream& stream);
    static void readXml(std::istream& stream, SerializationNode& node);
};

}  // namespace OpenMM
