// Registers the ATMMetaForce XML proxy when the library is loaded
// (reference: serialization/src/SerializationProxyRegistration.cpp:15-22).
#include <typeinfo>
#include "ATMMetaForce.h"
#include "ATMMetaForceProxy.h"
#include "openmm/serialization/SerializationProxy.h"

extern "C" __attribute__((visibility("default"))) void registerSerializationProxies() {
    static ATMMetaForcePlugin::ATMMetaForceProxy proxy;
    OpenMM::SerializationProxy::registerProxy(typeid(ATMMetaForcePlugin::ATMMetaForce), &proxy);
}

namespace {
struct AutoRegister {
    AutoRegister() { registerSerializationProxies(); }
} autoRegister;
}  // namespace
